/*
 * swap_hip.h -- C-ABI of the MI355X (gfx950) hybrid-PIC library.
 *
 * Each entry point replaces the body of one routine of the reference's
 * Fortran time step; the citation on each declaration names the reference
 * routine it stands in for.  The Fortran side keeps the reference's
 * subroutine names and argument lists and reaches these functions through
 * ISO_C_BINDING (swap_amd/fortran/swap_f.f90).
 *
 * Conventions shared by every function:
 *  - return 0 on success, a negative SWAP_E* code on failure; the text of
 *    the last failure is swap_last_error(ctx).  No C++ exception crosses
 *    this boundary.
 *  - work is enqueued on the context's own HIP stream; functions that hand
 *    data back to host memory (swap_get_*, swap_check_time_step, diagnostics)
 *    synchronise that stream before returning.
 *  - state is device resident.  Host arrays are mirrors moved only by the
 *    explicit swap_get_* / swap_set_* calls.
 *  - grid arrays use the reference's Fortran layouts: scalars (nx,ny,nz),
 *    vectors (nx,ny,nz,3), with nx = ncx+2 etc. (two ghost layers,
 *    dimensions.f:5).  Particle arrays are (n,3) column-major, i.e. the
 *    reference's (Ni_max,3) with Ni_max = n.
 *  - a context is not thread safe; separate contexts are independent.
 */
#ifndef SWAP_HIP_H
#define SWAP_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct swap_ctx swap_ctx;

enum {
  SWAP_OK = 0,
  SWAP_EINVAL = -1,   /* bad argument */
  SWAP_ENOMEM = -2,   /* device allocation failed */
  SWAP_EHIP = -3,     /* HIP runtime error */
  SWAP_ECAP = -4,     /* particle capacity exceeded */
  SWAP_ESTATE = -5    /* call out of order (e.g. no particles loaded) */
};

/* precision of particle state, moments and E (B is always advanced in FP64) */
enum { SWAP_FP32 = 32, SWAP_FP64 = 64 };

/* config flags */
enum {
  SWAP_SPECIES = 1,        /* per-particle beta_p and mrat arrays exist */
  SWAP_OUTFLOW_X = 2,      /* x is inflow/outflow instead of periodic
                              (build-defined; the reference is periodic,
                              gutsp.f:490-510, boundary.f:9-52): ghosts in x
                              are zero-gradient copies, the density of the
                              face nodes 1 and ncx is divided by their share
                              of the particle domain (1.5 and 0.5 cells),
                              ions that leave through either x face during
                              swap_step stay in the arrays as dead cells,
                              which every kernel skips, until the next
                              scheduled sort drops them (swap_diag.escaped
                              counts them then); entry points that hand out
                              particles, their count or particle
                              diagnostics drop them first (a sort), and the
                              sources do when the capacity would run out.
                              The capacity must hold the ions that left
                              since the last sort.  swap_set_inflow feeds
                              the x = 0 face.  Needs ncx >= 4 and
                              sort_every >= 1; no denoiser. */
  SWAP_TRACK_IDS = 4,      /* keep a particle id through sorts (tests) */
  SWAP_UNFUSED_FIELDS = 8, /* field subcycles of swap_step as one kernel per
                              reference routine (nine launches) instead of the
                              fused LDS-tiled predictor and corrector (two);
                              the same arithmetic, kept for checking */
  SWAP_PERSISTENT_FIELDS = 16 /* the fused subcycles of a step (or of one
                              swap_field_subcycles call) as one cooperative
                              launch with a grid barrier between the
                              predictor and corrector phases, instead of two
                              launches per subcycle; bit-identical, measured
                              slower, kept for comparison.  Ignored with
                              SWAP_UNFUSED_FIELDS. */
};

typedef struct {
  int ncx, ncy, ncz;       /* physical cells per axis (reference nx-2 ...) */
  double dx, dy, dz;       /* uniform cell size, km (inputs.f:136-138) */
  double dt;               /* particle step, s (inputs.f:144) */
  int ntsub;               /* field subcycles per step (inputs.f:27) */
  double alpha;            /* (mu0/1e3) q^2/m_ion (inputs.f:170) */
  double beta;             /* particles per physical density unit */
  double nu;               /* collision frequency (grd6_setup, initial.f:79) */
  int64_t capacity;        /* particle slots (reference Ni_max) */
  int precision;           /* SWAP_FP32 or SWAP_FP64 */
  int flags;               /* SWAP_SPECIES | SWAP_OUTFLOW_X | ... */
  int sort_every;          /* fused step: counting sort at least every k steps
                              (0: never); earlier when the last step left a
                              work item whose box outgrew its deposit window */
  uint64_t seed;           /* counter-based RNG key */
} swap_config;

/* field ids for swap_get_field / swap_set_field / swap_periodic / curls */
enum {
  SWAP_F_B1 = 0,    /* b1 == bt (f_update_tlev keeps them equal) */
  SWAP_F_B12 = 1,
  SWAP_F_B1P2 = 2,
  SWAP_F_BTC = 3,   /* node B, edge_to_center(bt) */
  SWAP_F_NP = 4,    /* scalar */
  SWAP_F_UP = 5,
  SWAP_F_AJ = 6,
  SWAP_F_E = 7,
  SWAP_F_NU = 8,    /* scalar */
  SWAP_F_AJC = 9,   /* node aj */
  SWAP_F_UPC = 10,  /* node up */
  SWAP_F_TMP0 = 11, /* general vector scratch fields for the routine-level */
  SWAP_F_TMP1 = 12, /* API (curlE output, face/edge interpolation results) */
  SWAP_F_TMP2 = 13,
  SWAP_F_COUNT = 14
};

/* particle arrays for swap_get_particle_array */
enum {
  SWAP_P_XP = 0,     /* absolute position (n,3), km, FP64 on the host side */
  SWAP_P_VP = 1,     /* (n,3) */
  SWAP_P_VP1 = 2,    /* (n,3) */
  SWAP_P_EP = 3,     /* (n,3)  get_Ep output */
  SWAP_P_VPLUS = 4,  /* (n,3) */
  SWAP_P_VMINUS = 5, /* (n,3) */
  SWAP_P_IJKP = 6,   /* (n,3) int32, 1-based cell indices */
  SWAP_P_WGHT = 7,   /* (n,8) */
  SWAP_P_CELL = 8,   /* (n) uint32 packed i | j<<11 | k<<21, 0-based cells */
  SWAP_P_FRAC = 9,   /* (n,3) in-cell fraction, (0,1] */
  SWAP_P_BETA = 10,  /* (n) beta_p */
  SWAP_P_MRAT = 11,  /* (n) mrat */
  SWAP_P_ID = 12     /* (n) uint32 particle id (SWAP_TRACK_IDS) */
};

/* update_up velocity sources */
enum { SWAP_V_CUR = 0, SWAP_V_EXTRAP = 1, SWAP_V_VP1 = 2 };

typedef struct {
  double charge;        /* sum 1/(beta beta_p) */
  double momentum[3];   /* sum m v /(beta beta_p), kg km/s (m = m_p/mrat) */
  double kinetic;       /* sum 1/2 m |v|^2 /(beta beta_p), J */
  double divb_l2;       /* sqrt(sum_cells (div b1)^2) */
  double b_energy;      /* sum_cells |b1|^2 (q/m_p units) */
  int64_t nparticles;
  int64_t escaped;      /* cumulative outflow removals */
  int64_t injected;     /* cumulative injections */
  int64_t window_misses;/* particles that took the global-atomic deposit path */
  int64_t shrunk_windows;/* P3 work items whose deposit window gave up part of
                            its shift range to fit LDS (their fastest
                            particles become window misses) */
  int64_t sorts;        /* counting sorts done by swap_sort / swap_step */
} swap_diag;

/* ---- lifecycle (replaces dimensions.f/global.f static state) ---- */
int swap_create(const swap_config *cfg, swap_ctx **out);
int swap_destroy(swap_ctx *ctx);
const char *swap_last_error(const swap_ctx *ctx);
int swap_sync(swap_ctx *ctx);
int64_t swap_num_particles(const swap_ctx *ctx);
/* the Philox key of later loads and injections (swap_config.seed at
   creation): a fresh seeded realisation in the same context */
int swap_set_seed(swap_ctx *ctx, uint64_t seed);
/* geometry: qx(i) etc. as the device uses them (grd7, initial.f:147-301) */
int swap_get_coords(swap_ctx *ctx, int axis, double *q /* n_axis+2 */);

/* ---- particles ---- */
/* reference layout: xp(n,3) absolute positions in (q(1), q(n-1)],
   vp(n,3), vp1(n,3); beta_p, mrat may be NULL (then 1).  Cell search follows
   get_interp_weights (gutsp.f:1391-1412): q(i) < x <= q(i+1). */
int swap_load_particles(swap_ctx *ctx, int64_t n, const double *xp, const double *vp,
                        const double *vp1, const double *beta_p, const double *mrat);
/* split form: packed cells and fractions as stored on the device; a cell
   outside the grid or a fraction outside (0, 1] is refused (SWAP_EINVAL) */
int swap_load_particles_split(swap_ctx *ctx, int64_t n, const uint32_t *cell,
                              const void *frac, const void *vp, const void *vp1,
                              const float *beta_p, const float *mrat);
/* seeded drifting-Maxwellian load, the sampler of load_Maxwellian
   (part_init.f:176-206) driven by Philox-4x32-10 keyed by (seed, id) */
int swap_load_maxwellian(swap_ctx *ctx, int64_t n, double vsw, double vth,
                         double beta_p, double mrat);
/* host copy of one particle array (see SWAP_P_*); dtype: FP64 for XP, the
   context precision for real arrays, uint32/int32 for integer arrays */
int swap_get_particle_array(swap_ctx *ctx, int which, void *host);
/* No counterpart in the reference: swap_get_particle_array for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving
   loops, hosts that forbid hipMalloc).  SWAP_P_XP is staged through a fixed
   device buffer of the library -- SWAP_NOALLOC_MAX particles, 24 MiB on each
   GPU it runs on, held until the process ends -- instead of a hipMalloc per
   call; a context with more particles is refused with SWAP_EINVAL.  Every
   call on one GPU shares that buffer: calls that overlap on one GPU (two host
   threads, two contexts) race.  Other arrays are copied as
   swap_get_particle_array copies them, which allocates nothing. */
#define SWAP_NOALLOC_MAX (1 << 20)
int swap_get_particle_array_noalloc(swap_ctx *ctx, int which, void *host);
/* host -> device for one array of the n loaded particles, with the layouts
   and dtypes of swap_get_particle_array.  XP re-splits the positions into
   cell + fraction.  EP, VPLUS, VMINUS, WGHT and IJKP fill the routine-level
   arrays, so a shim that keeps the reference's host arrays (xp, vp, vp1,
   Ep, vplus, vminus of gutsp.f) can pass them in before each routine. */
int swap_set_particle_array(swap_ctx *ctx, int which, const void *host);

/* ---- fields ---- */
int swap_set_field(swap_ctx *ctx, int field, const double *host);
int swap_get_field(swap_ctx *ctx, int field, double *host);
/* the same for REAL*4 host arrays, as the reference declares them */
int swap_set_field_f32(swap_ctx *ctx, int field, const float *host);
int swap_get_field_f32(swap_ctx *ctx, int field, float *host);
/* periodic / periodic_scalar (boundary.f:9-52, 167-206) */
int swap_periodic(swap_ctx *ctx, int field);

/* ---- routine-level API (one call per reference routine) ---- */
int swap_get_interp_weights(swap_ctx *ctx);        /* gutsp.f:1378 */
int swap_update_np(swap_ctx *ctx);                 /* gutsp.f:1441 */
int swap_update_up(swap_ctx *ctx, int vsrc);       /* gutsp.f:1649 */
int swap_extrapol_up(swap_ctx *ctx);               /* gutsp.f:127 */
int swap_get_Ep(swap_ctx *ctx);                    /* gutsp.f:155 */
int swap_get_vplus_vminus(swap_ctx *ctx);          /* gutsp.f:281 */
int swap_improve_up(swap_ctx *ctx);                /* gutsp.f:407 */
int swap_get_vp_final(swap_ctx *ctx);              /* gutsp.f:433 */
int swap_move_ion_half(swap_ctx *ctx);             /* gutsp.f:454 */
int swap_curlB(swap_ctx *ctx, int b_in, int aj_out);      /* gutsf.f:178 */
int swap_curlE(swap_ctx *ctx, int e_in, int out);         /* gutsf.f:249 */
int swap_crossf2(swap_ctx *ctx, int aa, int btc, int out);/* gutsf.f:40 */
int swap_get_E(swap_ctx *ctx);                     /* gutsf.f:994 */
int swap_predict_B(swap_ctx *ctx);                 /* gutsf.f:1115 */
int swap_get_Ep1(swap_ctx *ctx);                   /* gutsf.f:1174 */
int swap_correct_B(swap_ctx *ctx);                 /* gutsf.f:1309 */
int swap_f_update_tlev(swap_ctx *ctx);             /* gutsf.f:10 */
/* gutsf.f:1776; resets dtsub=dt/ntsub, ntf=ntsub first as maind.f:420-421
   does, returns the adjusted values */
int swap_check_time_step(swap_ctx *ctx, int *ntf, double *dtsub);
int swap_edge_to_center(swap_ctx *ctx, int in, int out);  /* grid_interp.f:10 */
int swap_edge_to_face(swap_ctx *ctx, int in, int out);    /* grid_interp.f:67 */
int swap_face_to_center(swap_ctx *ctx, int in, int out);  /* grid_interp.f:136 */

/* ---- fused time step (maind.f:385-479 as three particle passes) ---- */
/* curlB .. f_update_tlev for ntf subcycles (maind.f:425-476) */
int swap_field_subcycles(swap_ctx *ctx, int ntf, double dtsub);
/* counting sort of particles by tile/cell (build-only) */
int swap_sort(swap_ctx *ctx);
/* moments at the current positions: np and extrapolated up (the first
   get_interp_weights/update_np/extrapol_up of maind.f:385-399) */
int swap_prime(swap_ctx *ctx);
/* nsteps full steps; swap_prime must have run after the last load */
int swap_step(swap_ctx *ctx, int nsteps);

/* ---- sources (chem_rates.f) ---- */
/* Ionize_sw_mp, chem_rates.f:508-646: ring-distribution pickup protons,
   uniform in the domain, beta_p = beta_pu; returns the count injected */
int swap_ionize_sw(swap_ctx *ctx, int step, double ndot, double vsw, double beta_pu,
                   int64_t *injected);
/* spherical neutral cloud (chem_rates.f:290-503 form with the exp(-r/H)
   profile), photoionization + charge exchange (res_chex chem_rates.f:805) */
typedef struct {
  double cx, cy, cz;     /* cloud centre, km */
  double N0, r0, H;      /* n_n(r) = N0 (r0/r)^2 exp(-r/H), km^-3, km, km */
  double tau_photo;      /* s */
  double sigma_chex;     /* km^2 (chem_rates.f:817: 8e-26) */
  double mrat_pu;        /* 1/28 */
  double beta_min;       /* smallest beta_p of a new ion (in [0, 1]) */
  int max_per_cell;      /* a node with n > max_per_cell new ions gets
                            m = max(max_per_cell, ceil(beta_min n)) ions of
                            beta_p = m/n, so the deposit is unchanged */
  double rmin;           /* cloud core cut-off radius, km */
} swap_cloud;
int swap_ionize_cloud(swap_ctx *ctx, int step, const swap_cloud *cloud, int64_t *injected,
                      int64_t *chex);

/* Inflow through the x = 0 face of a SWAP_OUTFLOW_X context (build-defined;
   the reference's nearest routine refills a boundary layer to the upstream
   density, check_min_den_boundary gutsp.f:575-816).  Once set, every
   swap_step adds the ions of an upstream drifting Maxwellian (ppc ions per
   cell, bulk vsw along x, thermal vth as swap_load_maxwellian draws it)
   that cross the face during the step (the exact kinetic flux): candidates
   fill a slab of L = ceil((|vsw| + 4.08 vth) dt/dx) cells in front of the
   face at the start of the step, round(ppc L) in each y-z column of cells,
   and make the step's two half moves; one
   inside after the first deposits into the step's moments at x^{n+1/2},
   one inside after the second is appended with vp1 = vp, mrat = 1 and
   beta_p as given, its id continuing the id counter.  The step count keys
   the draws.  ppc = 0 switches it off. */
int swap_set_inflow(swap_ctx *ctx, double ppc, double vsw, double vth, double beta_p);

/* ---- diagnostics (FP64 reductions; Energy_diag part_init.f:14-139) ---- */
int swap_diagnostics(swap_ctx *ctx, swap_diag *out);

/* Energy_diag (part_init.f:14-139), the record maind.f:503-504 writes to
   c.energy.dat each step: energies in J with the reference's constants
   (q = 1.6e-19, mu0 = 3.14159 * 4e-7, epsilon = 8.85e-12, inputs.f:19,50-57),
   m_q = mion/q, cell volume dx dy dz.  Sums run over the unique nodes (the
   reference also sums its i = 1 ghost layer and a single y plane of its
   quasi-1-D grid) and weight every ion by 1/(beta beta_p) (the reference
   multiplies by beta_p there, part_init.f:69-70).  Evp: ions' kinetic
   energy (mass mion/mrat); Euf: 1/2 mion np |up|^2; EB1x/y/z: 1/(2 mu0)
   (m_q b1_c)^2; EB1: the same of b1 - b0 over all components; EE:
   epsilon/2 (m_q E km_to_m)^2; input_E: the kinetic energy of every ion
   loaded or added (sources, inflow) less that of ions that left through an
   open x (the reference's input_E, global.f:47); total_E = Evp + EE + EB1
   and norm_E = total_E / input_E (maind.f's "Normalized energy"). */
typedef struct {
  double input_E, Evp, Euf, EB1, EB1x, EB1y, EB1z, EE, total_E, norm_E;
} swap_energy;
int swap_energy_diag(swap_ctx *ctx, double mion, const double b0[3], swap_energy *out);
/* restore the input-energy account (J, for ions of mass mion) after a
   resume; loads reset it to the loaded ions' kinetic energy */
int swap_set_input_energy(swap_ctx *ctx, double mion, double input_E);

/* ---- learned moment denoiser (build-only; no reference counterpart) ----
   A 3-D convolutional net, 4 -> 32 -> 32 -> 32 -> 4 channels, 3x3x3 kernels
   with periodic padding, leaky ReLU between layers and a residual output, that maps
   the moments (np, up) of a low-ppc deposit to those of a high-ppc one.
   Convolutions run on MFMA with BF16 operands and FP32 accumulation; master
   weights are FP32 and trained with Adam by the library's own forward and
   backward kernels, on pairs the library generates from its own particles.
   Channels are normalised: np/n_scale - 1 and up/v_scale.  Parameters, flat
   FP32 in layer order: W[l][tap][c_in][c_out] (tap = dx+1 + 3(dy+1) + 9(dz+1))
   then b[l][c_out]. */
typedef struct {
  double n_scale;   /* density unit of the normalisation (e.g. n0) */
  double v_scale;   /* velocity unit of the normalisation, km/s */
  int max_pairs;    /* training pairs kept on the device */
  int stride;       /* low-ppc input: particles with id % stride == 0, weight x stride */
  uint64_t seed;    /* weight initialisation and pair draws */
} swap_dn_config;
int64_t swap_dn_num_params(void);
int swap_dn_create(swap_ctx *ctx, const swap_dn_config *cfg);
/* append a training pair from the current particles (needs SWAP_TRACK_IDS):
   input = the stride subsample's moments, target (mode):
     SWAP_DN_ALL       all particles' moments;
     SWAP_DN_DISJOINT  the moments of the other particles only (id % stride
                       != 0, weight x stride/(stride-1)), whose sampling noise
                       is independent of the input's;
     SWAP_DN_DISJOINT_STEP  the same target from the sums of the last
                       swap_step's final deposit minus the subsample's, with
                       no pass over the other particles; refused (SWAP_ESTATE)
                       if particles or moment sums changed since that step.
   The context's np/up fields are left as they were. */
enum { SWAP_DN_ALL = 0, SWAP_DN_DISJOINT = 1, SWAP_DN_DISJOINT_STEP = 2 };
int swap_dn_add_pair(swap_ctx *ctx, int mode);
int swap_dn_num_pairs(swap_ctx *ctx);
/* copy pair k out ([4][ncz][ncy][ncx] FP32 each, normalised), or append a
   pair given on the host (e.g. held-out pairs of a high-ppc run, for the
   gate of a low-ppc context) */
int swap_dn_get_pair(swap_ctx *ctx, int k, float *x, float *target);
int swap_dn_put_pair(swap_ctx *ctx, const float *x, const float *target);
/* iters Adam steps (learning rate lr) on pairs drawn from [0, ntrain); *loss
   = mean squared error (normalised units) of the last step */
int swap_dn_train(swap_ctx *ctx, int iters, double lr, int ntrain, double *loss);
/* mean squared error over pairs [p0, p1) of the raw input and of the net's
   output against the target.  Opens the gate of swap_dn_enable when the net
   is better on these pairs, closes it otherwise. */
int swap_dn_eval(swap_ctx *ctx, int p0, int p1, double *err_raw, double *err_net);
/* on = 1: swap_step replaces each deposit's np/up by the net's output; refused
   (SWAP_ESTATE) while the gate is closed, so the step keeps plain deposition */
int swap_dn_enable(swap_ctx *ctx, int on);
/* one forward + backward on a host input/target ([4][ncz][ncy][ncx] FP32,
   normalised): *loss and the parameter gradients (flat FP32) -- the check of
   the backward kernels against a CPU computation */
int swap_dn_grad(swap_ctx *ctx, const float *x, const float *target, double *loss,
                 float *grads);
/* the net's output for a host input ([4][ncz][ncy][ncx] FP32, normalised) */
int swap_dn_infer(swap_ctx *ctx, const float *x, float *y);
int swap_dn_get_params(swap_ctx *ctx, float *params);
int swap_dn_set_params(swap_ctx *ctx, const float *params);

/* ---- run state for a checkpoint (no reference counterpart: the reference
   restarts from restart.vars, maind.f:588-607, which holds its module
   state).  Everything swap_step reads beside the particles and the B levels:
   the step count (it keys the inflow draws), the sort schedule (particles
   sorted, steps since the sort, the work items with their cell boxes and the
   velocity bounds of the last pass, which fix how the FP32 run sums group
   particles), the id counter, the injected/escaped counts, the inflow
   parameters, the input-energy account and the counters of swap_diag.
   Restore after loading the same particles (swap_load_particles_split in
   the saved order, ids with swap_set_particle_array) into a context created
   with the same configuration: the resumed run then continues bit for bit.
   The blob is opaque host memory of swap_run_state_size bytes. ---- */
int swap_run_state_size(swap_ctx *ctx, int64_t *bytes);
int swap_save_run_state(swap_ctx *ctx, void *host);
int swap_restore_run_state(swap_ctx *ctx, const void *host);

/* ---- build-only instrumentation: HIP events on the context stream around
   each stage of swap_step; no reference counterpart (the reference times
   only the whole run, maind.f:648-654) ---- */
enum {
  SWAP_K_P1 = 0,      /* particle pass 1 */
  SWAP_K_P2 = 1,      /* particle pass 2 */
  SWAP_K_P3 = 2,      /* particle pass 3 */
  SWAP_K_SORT = 3,    /* counting sort */
  SWAP_K_FIELD = 4,   /* all field subcycles of a step */
  SWAP_K_GLUE = 5,    /* moment finalisation and node sources */
  SWAP_K_COUNT = 6
};
int swap_profile(swap_ctx *ctx, int enable);  /* enabling clears the totals */
/* total milliseconds and launch counts per stage since enabling */
int swap_profile_read(swap_ctx *ctx, double *ms, int64_t *count, int n);

#ifdef __cplusplus
}
#endif
#endif
