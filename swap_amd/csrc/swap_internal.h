// swap_internal.h -- context layout and kernel launchers (not part of the ABI).
#pragma once
#include <string>
#include <vector>
#include "../../include/swap_hip.h"
#include "swap_common.h"

// particle-pass tiling: particles are counting-sorted by (tile, cell in
// tile); one work item is a run of at most ITEM_MAX particles of one tile
constexpr int TILE_X = 4, TILE_Y = 8, TILE_Z = 8;
constexpr int TILE_CELLS = TILE_X * TILE_Y * TILE_Z;
constexpr int ITEM_MAX = 32768;
constexpr int TAIL_ITEM = 4096;  // particles per work item of the unsorted tail
constexpr int PASS_WG = 512;   // threads per particle-pass workgroup
// LDS deposit window, accumulator values: 81.6 KB, so that two workgroups
// share a CU's 160 KB; divisible by 3 and 4 (the moments of P2 and P1/P3)
constexpr int WIN_VALUES = 10200;
constexpr int BOX_SORT_VALUES = WIN_VALUES;  // box (nodes x 4) that calls for a sort
constexpr float WIN_SHIFT = 3.f;  // largest half-move cell shift a P3 window covers
constexpr int SREC = 8;  // values per node of the gather source S (6 used, 2 zero)

struct Item {
  int32_t start, end, tile, pad;
};

template <typename R>
struct PDev {
  uint32_t* cell;
  R *fx, *fy, *fz, *vx, *vy, *vz, *v1x, *v1y, *v1z;
  float* bp;     // beta_p (null: 1)
  float* mr;     // mrat (null: 1)
  uint32_t* id;  // particle id (null: not tracked)
};

template <typename R>
struct PCoef {
  R dt;        // particle step
  R s[3];      // (dt/2)/d_axis, for the split-precision half move
  int nc[3];   // cells per axis
  int outflow; // x inflow/outflow instead of periodic
};

struct swap_ctx {
  swap_config cfg;
  Geo g;
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  int prec = 32;
  size_t rsz = 4;
  double qn = 0;  // 1/(dx dy dz beta): deposit weight -> density
  // grid state
  double *b1 = nullptr, *b12 = nullptr, *b1p2 = nullptr;  // FP64 B advance
  bool b1p2_alias = true;  // after f_update_tlev, b1p2 equals b1
  void* fld[SWAP_F_COUNT] = {};  // real fields (B ids unused here)
  void *a = nullptr, *ct = nullptr, *btc0 = nullptr, *cE = nullptr;  // scratch
  void* S = nullptr;   // gather source: 6 components (aj-up at nodes, btc)
  void* macc = nullptr;  // moment accumulators [4][nn], AccOf<R>::T (8 bytes)
  double dtsub = 0;
  int ntf = 0;
  // particles
  int64_t n = 0, cap = 0;
  uint32_t *cell = nullptr, *cell2 = nullptr;
  void *p[9] = {}, *p2[9] = {};  // fx fy fz vx vy vz v1x v1y v1z
  float *bp = nullptr, *mr = nullptr, *bp2 = nullptr, *mr2 = nullptr;
  uint32_t *id = nullptr, *id2 = nullptr;
  // routine-level particle scratch (allocated on first use)
  void *Ep = nullptr, *vpl = nullptr, *vmi = nullptr, *wght = nullptr;
  int32_t* ijkp = nullptr;
  // sort and work items
  int ntile[3] = {0, 0, 0};
  int64_t ntiles = 0, nkeys = 0;
  uint32_t *scnt = nullptr, *sstart = nullptr, *scur = nullptr, *snrun = nullptr,
           *srstart = nullptr, *sblk = nullptr;
  uint2* sruns = nullptr;     // segment records of the sort (<= n)
  uint32_t* srbase = nullptr; // per sorted chunk position: source | segment << 12
  uint32_t *segkey = nullptr, *seginfo = nullptr, *segbase = nullptr;  // per chunk segment
  uint32_t* nsegc = nullptr;  // segments per chunk
  // signed velocity bounds (float keys): 6 for the whole grid, then 6 per
  // tile; vb and vb_next are the two halves of vb_mem and swap every step
  int *vb = nullptr, *vb_next = nullptr, *vb_mem = nullptr;
  bool vb_valid = false;
  int* d_zero = nullptr;
  Item* items = nullptr;
  int* boxes = nullptr;       // [max_items][6]: cell box lo[3], hi[3]
  int* boxes_next = nullptr;
  int* tboxes = nullptr;      // [tail items][6]: boxes of the unsorted tail's work items
  int* d_nitems = nullptr;
  int64_t max_items = 0;
  int64_t n_sorted = -1;      // particles covered by the item list (-1: none)
  int steps_since_sort = 0;
  // mutation count of the particles and the moment sums (macc); step_gen is
  // its value when swap_step finished, so mgen == step_gen means macc still
  // holds the sums of the last step's final deposit over today's particles
  uint64_t mgen = 0, step_gen = ~0ull;
  // cloud source scratch (allocated on first use): per interior node
  uint32_t *ccnt = nullptr, *coff = nullptr;
  float* cbp = nullptr;
  // counters
  // [0] window misses [1] escapes [2] scratch [3] chex [4] items whose next
  // box (reduced by the last P3) outgrows a P1 window [5] P3 windows that
  // gave up part of their shift range to fit the LDS budget
  unsigned long long* dcnt = nullptr;
  unsigned long long box_over = 0;  // dcnt[4], read back with check_time_step
  int64_t injected = 0, escaped = 0;
  int64_t next_id = 0;  // id of the next appended ion (ids stay unique when ions leave)
  // inflow through an open x face (swap_set_inflow): upstream ions per cell,
  // drift, thermal speed, beta_p, slab depth (cells), slots per step, the
  // last step's count and the slot buffers
  double in_ppc = 0, in_vsw = 0, in_vth = 0, in_bp = 1;
  int in_L = 0;
  int64_t in_slots = 0, in_last = 0;
  uint32_t *icnt = nullptr, *ioff = nullptr, *iblk = nullptr;
  int64_t islots = 0;
  int fz_pred = 0, fz_corr = 0;  // z chunks of the fused field launches (0: not yet chosen)
  int pers_grid = 0, pers_tz = 0;  // cooperative subcycle launch: workgroups, planes per tile
  double* dred = nullptr;     // FP64 reduction scratch
  double* ts = nullptr;       // device dtsub and ntf of the current step (check_time_step)
  // pinned host copies of ts (dtsub, ntf) and of the box-overflow count, and
  // the event after which they hold the current step's values
  struct StepStat {
    double dtsub, ntf;
    unsigned long long box_over;
  }* hstat = nullptr;
  hipEvent_t ev_stat = nullptr;
  float* dmin = nullptr;      // check_time_step minimum
  void* scratch = nullptr;    // general device scratch
  size_t scratch_bytes = 0;
  int64_t step_count = 0;
  int64_t sorts = 0;          // counting sorts done
  // an open x: ions may have left since the last sort (cell DEAD_CELL, still
  // in the arrays until a sort drops them)
  bool dead_pending = false;
  // learned moment denoiser (k_denoise.hip); null until swap_dn_create
  struct DnState* dn = nullptr;
  // stage timing (swap_profile): event pairs recorded on `stream`
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev[SWAP_K_COUNT];  // start, stop, start, stop ...
};

// ints in each half of vb_mem: the grid's 6 bounds, then 6 per tile
inline int64_t vb_ints(const swap_ctx* c) { return 6 * (c->ntiles + 1); }

// error helpers
int set_err(swap_ctx* c, int code, const char* fmt, ...);
#define HIPCHK(c, call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_err((c), SWAP_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// ---- field launchers (k_fields.hip) ----
int fl_curlB(swap_ctx* c, const double* b, const double* b2, void* aj, void* a);
int fl_cross_nodes(swap_ctx* c, const void* a, const double* b, void* ct);
int fl_e2c_B(swap_ctx* c, const double* b, const double* b2, void* out);
int fl_cross_smooth(swap_ctx* c, const void* a, const void* btc0, void* ct);
int fl_edgeavg_E(swap_ctx* c, const void* ct, const void* aj, void* E);
int fl_curlE_update(swap_ctx* c, const void* E, const double* base, double coef, double* out);
int fl_curlE(swap_ctx* c, const void* E, void* out);
int fl_f2c(swap_ctx* c, const void* in, void* out);
int fl_e2c(swap_ctx* c, const void* in, void* out);
int fl_e2f(swap_ctx* c, const void* in, const double* bin, void* out);
int fl_crossf2(swap_ctx* c, const void* aa, const void* btc, void* out);
int fl_periodic(swap_ctx* c, void* f, int ncomp, int is_double);
int fl_a_from(swap_ctx* c, const void* aj, const void* up, void* a);
int fl_build_src(swap_ctx* c, const void* aj, const void* up, const void* btc, void* ajc,
                 void* S);
int fl_update_src(swap_ctx* c, const void* up, const void* ajc, void* S);
int fl_finalize_np(swap_ctx* c);
int fl_finalize_up(swap_ctx* c);
int fl_check_time_step(swap_ctx* c, int* ntf, double* dtsub);
// check_time_step into c->ts on the device (no host synchronisation)
int fl_time_step_dev(swap_ctx* c);
int fl_set_dtsub(swap_ctx* c, double dtsub, int ntf);
int fl_copy_b(swap_ctx* c, const double* src, double* dst);
int fl_subcycle_fused(swap_ctx* c, bool last);  // dtsub from c->ts
int fl_subcycles_persistent(swap_ctx* c, int nsub);  // nsub of them, one launch
int fl_zero(swap_ctx* c, void* p, size_t bytes);

// ---- particle launchers (k_particles.hip) ----
int pl_weights(swap_ctx* c);
// 0: np, 1: up(vp) 2: up(ext) 3: up(vp1); stride > 1: the id % stride == 0 subsample
int pl_deposit_simple(swap_ctx* c, int mode, int stride = 1, bool back = false);
int pl_get_Ep(swap_ctx* c);
int pl_vplus_vminus(swap_ctx* c);
int pl_improve_vp1(swap_ctx* c);
int pl_vp_final(swap_ctx* c);
int pl_move_half(swap_ctx* c);
int pl_pass(swap_ctx* c, int mode);  // fused passes 1, 2, 3
int pl_load_abs(swap_ctx* c, int64_t n, const double* xp_d, const double* v_d,
                const double* v1_d);
int pl_export_abs(swap_ctx* c, double* xp_d);
int pl_set_abs(swap_ctx* c, const double* xp_d);  // positions only
int pl_ensure_scratch(swap_ctx* c);               // Ep, vplus, vminus, wght, ijkp
int pl_maxwellian_full(swap_ctx* c, int64_t off, int64_t n, double vsw, double vth, float bpv,
                       float mrv, uint32_t stream);
int pl_box_reset(swap_ctx* c);

// ---- no-alloc staging (swap_noalloc.hip): device address of its buffer ----
int na_setup(swap_ctx* c);
// an open x: drop the ions that left since the last sort (swap_api.hip)
int compact(swap_ctx* c);

// ---- sort (k_sort.hip) ----
int so_sort(swap_ctx* c);
// exclusive scan of n uint32 counts into out[0..n], out[n] = total; bsum: n / 4096 + 1
// block sums, the context's own (n <= nkeys + 1) when null
int so_scan_u32(swap_ctx* c, int64_t n, const uint32_t* in, uint32_t* out,
                uint32_t* bsum = nullptr);

// ---- sources / diagnostics (k_sources.hip) ----
int sr_ionize_sw(swap_ctx* c, int step, double ndot, double vsw, double beta_pu,
                 int64_t* injected);
int sr_ionize_cloud(swap_ctx* c, int step, const swap_cloud* cl, int64_t* injected,
                    int64_t* chex);
int sr_diagnostics(swap_ctx* c, swap_diag* d);
// slots of the FP64 reduction scratch dred: the input-energy account and the
// energy diagnostic's sums
constexpr int DRED_INPUT = 32, DRED_ENERGY = 40;
int sr_input_add(swap_ctx* c, int64_t off, int64_t n, double sign);
int sr_input_reset(swap_ctx* c);
int sr_energy_diag(swap_ctx* c, double mion, const double b0[3], swap_energy* d);
int sr_set_inflow(swap_ctx* c, double ppc, double vsw, double vth, double beta_p);
int sr_inflow_step(swap_ctx* c);  // after the final particle pass of swap_step
int sr_vbounds(swap_ctx* c);  // signed per-axis velocity bounds into c->vb

// ---- denoiser (k_denoise.hip) ----
// replace np (with_np) and up by the net's output when enabled; no-op otherwise
int dn_apply_step(swap_ctx* c, bool with_np);
void dn_destroy(swap_ctx* c);

// helpers
template <typename R>
PDev<R> pdev(swap_ctx* c) {
  PDev<R> d;
  d.cell = c->cell;
  R** q = (R**)c->p;
  d.fx = q[0]; d.fy = q[1]; d.fz = q[2];
  d.vx = q[3]; d.vy = q[4]; d.vz = q[5];
  d.v1x = q[6]; d.v1y = q[7]; d.v1z = q[8];
  d.bp = c->bp;
  d.mr = c->mr;
  d.id = c->id;
  return d;
}
template <typename R>
PCoef<R> pcoef(swap_ctx* c) {
  PCoef<R> k;
  k.dt = (R)c->cfg.dt;
  k.s[0] = (R)(0.5 * c->cfg.dt / c->cfg.dx);
  k.s[1] = (R)(0.5 * c->cfg.dt / c->cfg.dy);
  k.s[2] = (R)(0.5 * c->cfg.dt / c->cfg.dz);
  k.nc[0] = c->cfg.ncx; k.nc[1] = c->cfg.ncy; k.nc[2] = c->cfg.ncz;
  k.outflow = (c->cfg.flags & SWAP_OUTFLOW_X) ? 1 : 0;
  return k;
}
inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }
