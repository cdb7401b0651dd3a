// swap_api.hip -- the C-ABI of include/swap_hip.h: context, state transfer,
// the routine-level entry points and the fused time step.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>
#include "swap_internal.h"

int set_err(swap_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

namespace {

bool is_b(int f) { return f == SWAP_F_B1 || f == SWAP_F_B12 || f == SWAP_F_B1P2; }
int ncomp_of(int f) { return (f == SWAP_F_NP || f == SWAP_F_NU) ? 1 : 3; }

double* bptr(swap_ctx* c, int f) {
  if (f == SWAP_F_B1) return c->b1;
  if (f == SWAP_F_B12) return c->b12;
  return c->b1p2_alias ? c->b1 : c->b1p2;
}

int dmalloc(swap_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess)
    return set_err(c, SWAP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  e = hipMemset(*p, 0, bytes ? bytes : 16);
  if (e != hipSuccess) return set_err(c, SWAP_EHIP, "hipMemset: %s", hipGetErrorString(e));
  return 0;
}
#define DM(ptr, bytes)                                          \
  do {                                                          \
    if (int e_ = dmalloc(c, (void**)&(ptr), (bytes))) return e_; \
  } while (0)

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// copy a real device array to a double/float host array, or back
int real_to_host(swap_ctx* c, const void* d, size_t n, double* hd, float* hf) {
  if (c->prec == 64 && hd) {
    HIPCHK(c, hipMemcpyAsync(hd, d, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  if (c->prec == 32 && hf) {
    HIPCHK(c, hipMemcpyAsync(hf, d, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  if (c->prec == 64) {
    std::vector<double> t(n);
    HIPCHK(c, hipMemcpyAsync(t.data(), d, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < n; q++) hf[q] = (float)t[q];
  } else {
    std::vector<float> t(n);
    HIPCHK(c, hipMemcpyAsync(t.data(), d, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < n; q++) hd[q] = t[q];
  }
  return 0;
}
int host_to_real(swap_ctx* c, void* d, size_t n, const double* hd, const float* hf) {
  if (c->prec == 64 && hd) {
    HIPCHK(c, hipMemcpyAsync(d, hd, n * 8, hipMemcpyHostToDevice, c->stream));
  } else if (c->prec == 32 && hf) {
    HIPCHK(c, hipMemcpyAsync(d, hf, n * 4, hipMemcpyHostToDevice, c->stream));
  } else if (c->prec == 64) {
    std::vector<double> t(hf, hf + n);
    HIPCHK(c, hipMemcpyAsync(d, t.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  } else {
    std::vector<float> t(n);
    for (size_t q = 0; q < n; q++) t[q] = (float)hd[q];
    HIPCHK(c, hipMemcpyAsync(d, t.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int zero_moments(swap_ctx* c, int m0, int m1) {
  c->mgen++;
  size_t off = 8 * (size_t)c->g.nn * m0, len = 8 * (size_t)c->g.nn * (m1 - m0);
  HIPCHK(c, hipMemsetAsync((char*)c->macc + off, 0, len, c->stream));
  return 0;
}

#define TRY(x)                  \
  do {                          \
    if (int e_ = (x)) return e_; \
  } while (0)

// stage timing: record an event on the context stream at a stage boundary
struct Stage {
  swap_ctx* c;
  int k;
  Stage(swap_ctx* cc, int kk) : c(cc), k(kk) { mark(); }
  ~Stage() { mark(); }
  void mark() {
    if (!c->prof_on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->prof_ev[k].push_back(e);
  }
};

void prof_clear(swap_ctx* c) {
  for (int k = 0; k < SWAP_K_COUNT; k++) {
    for (hipEvent_t e : c->prof_ev[k]) (void)hipEventDestroy(e);
    c->prof_ev[k].clear();
  }
}

int get_E_impl(swap_ctx* c) {
  TRY(fl_a_from(c, c->fld[SWAP_F_AJ], c->fld[SWAP_F_UP], c->a));
  TRY(fl_cross_nodes(c, c->a, c->b1, c->ct));
  return fl_edgeavg_E(c, c->ct, c->fld[SWAP_F_AJ], c->fld[SWAP_F_E]);
}
void materialise_b1p2(swap_ctx* c) {
  if (c->b1p2_alias) {
    (void)fl_copy_b(c, c->b1, c->b1p2);
    c->b1p2_alias = false;
  }
}
int get_Ep1_impl(swap_ctx* c) {
  materialise_b1p2(c);
  TRY(fl_curlB(c, c->b1p2, c->b1, c->fld[SWAP_F_AJ], c->a));
  TRY(fl_e2c_B(c, c->b1p2, c->b1, c->btc0));
  TRY(fl_cross_smooth(c, c->a, c->btc0, c->ct));
  return fl_edgeavg_E(c, c->ct, c->fld[SWAP_F_AJ], c->fld[SWAP_F_E]);
}
int tlev_impl(swap_ctx* c) {
  if (c->b1p2_alias) return 0;  // b1p2 already equals b1 and b12 update is b12 = b1
  double* old12 = c->b12;
  c->b12 = c->b1;
  c->b1 = c->b1p2;
  c->b1p2 = old12;
  c->b1p2_alias = true;
  return 0;
}
int subcycle_impl(swap_ctx* c, double dtsub) {
  // curlB(bt, np, aj) (maind.f:430) with a = aj - up for get_E (gutsf.f:1058)
  TRY(fl_curlB(c, c->b1, nullptr, c->fld[SWAP_F_AJ], c->a));
  // predict_B (gutsf.f:1115-1169)
  TRY(fl_cross_nodes(c, c->a, c->b1, c->ct));
  TRY(fl_edgeavg_E(c, c->ct, c->fld[SWAP_F_AJ], c->fld[SWAP_F_E]));
  TRY(fl_curlE_update(c, c->fld[SWAP_F_E], c->b12, 2.0 * dtsub, c->b1p2));
  c->b1p2_alias = false;
  // correct_B (gutsf.f:1309-1359)
  TRY(get_Ep1_impl(c));
  TRY(fl_curlE_update(c, c->fld[SWAP_F_E], c->b1, dtsub, c->b1p2));
  // f_update_tlev (gutsf.f:10-35)
  return tlev_impl(c);
}
// the same subcycle as two LDS-tiled launches; the corrector writes into
// b12's buffer, so f_update_tlev is the rotation b12 <- b1 <- corrected
// a load starts the input-energy account from the loaded ions
int load_input_energy(swap_ctx* c) {
  TRY(sr_input_reset(c));
  return sr_input_add(c, 0, c->n, 1.0);
}

int subcycle_fused(swap_ctx* c, bool last) {
  TRY(fl_subcycle_fused(c, last));
  double* corrected = c->b12;
  c->b12 = c->b1;
  c->b1 = corrected;
  c->b1p2_alias = true;
  return 0;
}

// nsub fused subcycles, the last storing E and aj: one cooperative launch
// with SWAP_PERSISTENT_FIELDS, else two launches each; the B buffers end as
// nsub calls of subcycle_fused leave them
int subcycles_fused(swap_ctx* c, int nsub) {
  if (!(c->cfg.flags & SWAP_PERSISTENT_FIELDS)) {
    for (int q = 0; q < nsub; q++) TRY(subcycle_fused(c, q == nsub - 1));
    return 0;
  }
  if (nsub < 1) return 0;
  TRY(fl_subcycles_persistent(c, nsub));
  if (nsub & 1) std::swap(c->b1, c->b12);
  c->b1p2_alias = true;
  return 0;
}

}  // namespace

// An open x keeps the ions that left as dead cells until the next sort;
// what hands out particles, their count or particle diagnostics drops them
// first (a sort), so callers see the live ions only.
int compact(swap_ctx* c) {
  if (!c->g.ox || !c->dead_pending) return 0;
  return swap_sort(c);
}

extern "C" {

int swap_create(const swap_config* cfg, swap_ctx** out) {
  if (!cfg || !out) return SWAP_EINVAL;
  *out = nullptr;
  swap_ctx* c = new (std::nothrow) swap_ctx();
  if (!c) return SWAP_ENOMEM;
  c->cfg = *cfg;
  auto fail = [&](int code) {
    swap_destroy(c);
    return code;
  };
  if (cfg->ncx < 2 || cfg->ncy < 2 || cfg->ncz < 2 || cfg->ncx > 2046 || cfg->ncy > 1022 ||
      cfg->ncz > 2046 || !(cfg->dx > 0) || !(cfg->dy > 0) || !(cfg->dz > 0) || !(cfg->dt > 0) ||
      cfg->ntsub < 1 || cfg->capacity < 0 || cfg->capacity >= (1ll << 31) ||
      (cfg->precision != 32 && cfg->precision != 64) || !(cfg->beta > 0) || !(cfg->alpha > 0)) {
    fprintf(stderr, "swap_create: invalid configuration\n");
    delete c;
    return SWAP_EINVAL;
  }
  {
    // the sort's chunk-relative key holds the tile offsets of a chunk that
    // spans the grid in 24 bits above the cell index (with one more tile for
    // the ions that left an open x): at most 23 bits of tile coordinates
    auto bits = [](int v) { return v > 0 ? 32 - __builtin_clz((unsigned)v) : 0; };
    if (cfg->ncx >= 2 && cfg->ncy >= 2 && cfg->ncz >= 2 &&
        bits((cfg->ncx + TILE_X - 1) / TILE_X - 1) + bits((cfg->ncy + TILE_Y - 1) / TILE_Y - 1) +
                bits((cfg->ncz + TILE_Z - 1) / TILE_Z - 1) > 23) {
      fprintf(stderr, "swap_create: grid too large for the sort keys (at most 23 bits of "
                      "tile coordinates, e.g. 2046 x 1022 x 1024 cells)\n");
      delete c;
      return SWAP_EINVAL;
    }
  }
  if ((cfg->flags & SWAP_OUTFLOW_X) && (cfg->ncx < 4 || cfg->sort_every < 1)) {
    fprintf(stderr, "swap_create: an open x needs ncx >= 4 and sort_every >= 1 "
                    "(the sort removes escaped ions)\n");
    delete c;
    return SWAP_EINVAL;
  }
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    fprintf(stderr, "swap_create: no HIP device\n");
    delete c;
    return SWAP_EHIP;
  }
  c->prec = cfg->precision;
  c->rsz = c->prec == 64 ? 8 : 4;
  Geo& g = c->g;
  g.ncx = cfg->ncx; g.ncy = cfg->ncy; g.ncz = cfg->ncz;
  g.nx = g.ncx + 2; g.ny = g.ncy + 2; g.nz = g.ncz + 2;
  g.nn = (int64_t)g.nx * g.ny * g.nz;
  g.ox = (cfg->flags & SWAP_OUTFLOW_X) ? 1 : 0;
  c->qn = 1.0 / (cfg->dx * cfg->dy * cfg->dz * cfg->beta);
  c->dtsub = cfg->dt / cfg->ntsub;
  c->ntf = cfg->ntsub;
  size_t nn = (size_t)g.nn, rs = c->rsz;
  int e = 0;
#define DMF(ptr, bytes)                               \
  do {                                                \
    if ((e = dmalloc(c, (void**)&(ptr), (bytes)))) {  \
      fprintf(stderr, "swap_create: %s\n", c->err.c_str()); \
      return fail(e);                                 \
    }                                                 \
  } while (0)
  DMF(c->b1, 24 * nn);
  DMF(c->b12, 24 * nn);
  DMF(c->b1p2, 24 * nn);
  for (int f = 0; f < SWAP_F_COUNT; f++) {
    if (is_b(f)) continue;
    DMF(c->fld[f], rs * ncomp_of(f) * nn);
  }
  DMF(c->a, rs * 3 * nn);
  DMF(c->ct, rs * 3 * nn);
  DMF(c->btc0, rs * 3 * nn);
  DMF(c->S, rs * SREC * nn);  // node-major gather source
  DMF(c->macc, 8 * 4 * nn);
  c->scratch_bytes = std::max<size_t>(8 * nn, 1 << 16);
  DMF(c->scratch, c->scratch_bytes);
  DMF(c->dred, 64 * sizeof(double));
  DMF(c->ts, 2 * sizeof(double));
  if (hipHostMalloc((void**)&c->hstat, sizeof(*c->hstat), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_stat, hipEventDisableTiming) != hipSuccess) {
    fprintf(stderr, "swap_create: pinned step status\n");
    return fail(SWAP_ENOMEM);
  }
  DMF(c->dmin, 16 * sizeof(float));
  DMF(c->dcnt, 8 * sizeof(unsigned long long));
  DMF(c->d_nitems, 4 * sizeof(int));
  // particles (double-buffered for the sort)
  c->cap = cfg->capacity;
  size_t cp = (size_t)c->cap;
  DMF(c->cell, 4 * cp);
  DMF(c->cell2, 4 * cp);
  for (int a = 0; a < 9; a++) {
    DMF(c->p[a], rs * cp);
    DMF(c->p2[a], rs * cp);
  }
  if (cfg->flags & SWAP_SPECIES) {
    DMF(c->bp, 4 * cp);
    DMF(c->bp2, 4 * cp);
    DMF(c->mr, 4 * cp);
    DMF(c->mr2, 4 * cp);
  }
  if (cfg->flags & SWAP_TRACK_IDS) {
    DMF(c->id, 4 * cp);
    DMF(c->id2, 4 * cp);
  }
  // sort keys and work items
  c->ntile[0] = (g.ncx + TILE_X - 1) / TILE_X;
  c->ntile[1] = (g.ncy + TILE_Y - 1) / TILE_Y;
  c->ntile[2] = (g.ncz + TILE_Z - 1) / TILE_Z;
  c->ntiles = (int64_t)c->ntile[0] * c->ntile[1] * c->ntile[2];
  c->nkeys = c->ntiles * TILE_CELLS;
  DMF(c->sruns, 8 * cp);
  DMF(c->srbase, 4 * cp);
  DMF(c->segkey, 4 * cp);
  DMF(c->seginfo, 4 * cp);
  DMF(c->segbase, 4 * cp);
  DMF(c->nsegc, 4 * (cp / 4096 + 2));
  // counts and starts per key, plus the key of escaped particles and the total
  DMF(c->scnt, 4 * (size_t)(c->nkeys + 2));
  DMF(c->sstart, 4 * (size_t)(c->nkeys + 2));
  DMF(c->scur, 4 * (size_t)(c->nkeys + 2));
  DMF(c->snrun, 4 * (size_t)(c->nkeys + 2));
  DMF(c->srstart, 4 * (size_t)(c->nkeys + 2));
  DMF(c->vb_mem, 2 * vb_ints(c) * sizeof(int));
  c->vb = c->vb_mem;
  c->vb_next = c->vb_mem + vb_ints(c);
  DMF(c->d_zero, 4 * sizeof(int));
  DMF(c->sblk, 4 * (size_t)(c->nkeys / 4096 + 4));
  c->max_items = c->ntiles + c->cap / ITEM_MAX + 1;
  DMF(c->items, sizeof(Item) * (size_t)c->max_items);
  DMF(c->boxes, 24 * (size_t)c->max_items);
  DMF(c->boxes_next, 24 * (size_t)c->max_items);
  DMF(c->tboxes, 24 * (size_t)(c->cap / TAIL_ITEM + 2));
#undef DMF
  // the fixed staging buffer of swap_get_particle_array_noalloc on this device
  if ((e = na_setup(c))) {
    fprintf(stderr, "swap_create: %s\n", c->err.c_str());
    return fail(e);
  }
  // nu = nu_init everywhere (grd6_setup, initial.f:79)
  std::vector<double> nuh(nn, cfg->nu);
  if ((e = host_to_real(c, c->fld[SWAP_F_NU], nn, nuh.data(), nullptr))) return fail(e);
  *out = c;
  return 0;
}

int swap_destroy(swap_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  prof_clear(c);
  dn_destroy(c);
  void* ps[] = {c->b1, c->b12, c->b1p2, c->a, c->ct, c->btc0, c->cE, c->S, c->macc,
                c->scratch, c->dred, c->dmin, c->dcnt, c->d_nitems, c->cell, c->cell2, c->bp,
                c->bp2, c->mr, c->mr2, c->id, c->id2, c->Ep, c->vpl, c->vmi, c->wght, c->ijkp,
                c->sruns, c->srbase, c->segkey, c->seginfo, c->segbase, c->nsegc, c->scnt,
                c->sstart, c->scur, c->snrun, c->srstart, c->sblk, c->items, c->ccnt,
                c->coff, c->cbp, c->icnt, c->ioff, c->iblk,
                c->boxes, c->boxes_next, c->tboxes, c->vb_mem, c->d_zero};
  for (void* p : ps) dfree(p);
  dfree(c->ts);
  if (c->hstat) (void)hipHostFree(c->hstat);
  if (c->ev_stat) (void)hipEventDestroy(c->ev_stat);
  for (int f = 0; f < SWAP_F_COUNT; f++) dfree(c->fld[f]);
  for (int a = 0; a < 9; a++) {
    dfree(c->p[a]);
    dfree(c->p2[a]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char* swap_last_error(const swap_ctx* c) { return c ? c->err.c_str() : "null context"; }

int swap_sync(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swap_set_seed(swap_ctx* c, uint64_t seed) {
  if (!c) return SWAP_EINVAL;
  c->cfg.seed = seed;
  return 0;
}

int64_t swap_num_particles(const swap_ctx* cc) {
  if (!cc) return -1;
  swap_ctx* c = const_cast<swap_ctx*>(cc);  // dropping the ions that left changes no state a caller sees
  if (compact(c)) return -1;
  return c->n;
}

int swap_get_coords(swap_ctx* c, int axis, double* q) {
  if (!c || !q || axis < 0 || axis > 2) return SWAP_EINVAL;
  int n = axis == 0 ? c->g.nx : (axis == 1 ? c->g.ny : c->g.nz);
  double d = axis == 0 ? c->cfg.dx : (axis == 1 ? c->cfg.dy : c->cfg.dz);
  for (int i = 0; i < n; i++) q[i] = (double)(i + 1) * d;  // q(i) = i d, grd7 with no stretch
  return 0;
}

// ---------------- particles ----------------

int swap_load_particles(swap_ctx* c, int64_t n, const double* xp, const double* vp,
                        const double* vp1, const double* beta_p, const double* mrat) {
  if (!c || n < 0 || (n > 0 && (!xp || !vp || !vp1))) return SWAP_EINVAL;
  c->mgen++;
  if (n > c->cap) return set_err(c, SWAP_ECAP, "load: %lld > capacity", (long long)n);
  if ((beta_p || mrat) && !(c->cfg.flags & SWAP_SPECIES))
    return set_err(c, SWAP_EINVAL, "load: beta_p/mrat need SWAP_SPECIES");
  size_t b = sizeof(double) * 3 * (size_t)n;
  double* dtmp = nullptr;
  if (n > 0) {
    HIPCHK(c, hipMalloc(&dtmp, 3 * b));
    HIPCHK(c, hipMemcpyAsync(dtmp, xp, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((char*)dtmp + b, vp, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((char*)dtmp + 2 * b, vp1, b, hipMemcpyHostToDevice, c->stream));
  }
  c->n = n;
  c->next_id = n;
  int e = pl_load_abs(c, n, dtmp, (double*)((char*)dtmp + b), (double*)((char*)dtmp + 2 * b));
  if (!e && (c->cfg.flags & SWAP_SPECIES)) {
    std::vector<float> t(std::max<int64_t>(n, 1));
    for (int64_t q = 0; q < n; q++) t[q] = beta_p ? (float)beta_p[q] : 1.0f;
    if (n) (void)hipMemcpyAsync(c->bp, t.data(), 4 * n, hipMemcpyHostToDevice, c->stream);
    (void)hipStreamSynchronize(c->stream);
    for (int64_t q = 0; q < n; q++) t[q] = mrat ? (float)mrat[q] : 1.0f;
    if (n) (void)hipMemcpyAsync(c->mr, t.data(), 4 * n, hipMemcpyHostToDevice, c->stream);
  }
  if (!e && c->id) {
    std::vector<uint32_t> t(std::max<int64_t>(n, 1));
    for (int64_t q = 0; q < n; q++) t[q] = (uint32_t)q;
    if (n) (void)hipMemcpyAsync(c->id, t.data(), 4 * n, hipMemcpyHostToDevice, c->stream);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (dtmp) HIPCHK(c, hipFree(dtmp));
  c->n_sorted = -1;
  c->vb_valid = false;
  if (!e) TRY(load_input_energy(c));
  return e;
}

int swap_load_particles_split(swap_ctx* c, int64_t n, const uint32_t* cell, const void* frac,
                              const void* vp, const void* vp1, const float* beta_p,
                              const float* mrat) {
  if (!c || n < 0 || (n > 0 && (!cell || !frac || !vp || !vp1))) return SWAP_EINVAL;
  c->mgen++;
  if (n > c->cap) return set_err(c, SWAP_ECAP, "load: %lld > capacity", (long long)n);
  if ((beta_p || mrat) && !(c->cfg.flags & SWAP_SPECIES))
    return set_err(c, SWAP_EINVAL, "load: beta_p/mrat need SWAP_SPECIES");
  size_t rn = c->rsz * (size_t)n;
  for (int64_t q = 0; q < n; q++) {
    uint32_t v = cell[q];
    if (cell_i(v) >= c->cfg.ncx || cell_j(v) >= c->cfg.ncy || cell_k(v) >= c->cfg.ncz)
      return set_err(c, SWAP_EINVAL, "load: particle %lld has a cell outside the grid",
                     (long long)q);
  }
  // fractions lie in (0, 1] (right-closed cells); anything else would give
  // negative or extrapolated weights
  for (int64_t q = 0; q < 3 * n; q++) {
    const double f = c->prec == 64 ? ((const double*)frac)[q] : (double)((const float*)frac)[q];
    if (!(f > 0.0 && f <= 1.0))
      return set_err(c, SWAP_EINVAL, "load: particle %lld has a fraction %g outside (0, 1]",
                     (long long)(q % n), f);
  }
  c->n = n;
  c->next_id = n;
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(c->cell, cell, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    for (int a = 0; a < 3; a++) {
      HIPCHK(c, hipMemcpyAsync(c->p[a], (const char*)frac + a * rn, rn, hipMemcpyHostToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->p[3 + a], (const char*)vp + a * rn, rn, hipMemcpyHostToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->p[6 + a], (const char*)vp1 + a * rn, rn, hipMemcpyHostToDevice,
                               c->stream));
    }
    if (c->cfg.flags & SWAP_SPECIES) {
      std::vector<float> one(n, 1.0f);
      HIPCHK(c, hipMemcpyAsync(c->bp, beta_p ? beta_p : one.data(), 4 * n, hipMemcpyHostToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->mr, mrat ? mrat : one.data(), 4 * n, hipMemcpyHostToDevice,
                               c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (c->id) {
      std::vector<uint32_t> t(n);
      for (int64_t q = 0; q < n; q++) t[q] = (uint32_t)q;
      HIPCHK(c, hipMemcpyAsync(c->id, t.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_sorted = -1;
  c->vb_valid = false;
  return load_input_energy(c);
}

int swap_load_maxwellian(swap_ctx* c, int64_t n, double vsw, double vth, double beta_p,
                         double mrat) {
  if (!c || n < 0) return SWAP_EINVAL;
  c->mgen++;
  if (n > c->cap) return set_err(c, SWAP_ECAP, "load: %lld > capacity", (long long)n);
  if ((beta_p != 1.0 || mrat != 1.0) && !(c->cfg.flags & SWAP_SPECIES))
    return set_err(c, SWAP_EINVAL, "load: beta_p/mrat != 1 need SWAP_SPECIES");
  c->n = n;
  c->next_id = n;
  TRY(pl_maxwellian_full(c, 0, n, vsw, vth, (float)beta_p, (float)mrat, 1u));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_sorted = -1;
  c->vb_valid = false;
  return load_input_energy(c);
}

int swap_get_particle_array(swap_ctx* c, int which, void* host) {
  if (!c || !host) return SWAP_EINVAL;
  TRY(compact(c));
  size_t n = (size_t)c->n, rn = c->rsz * n;
  switch (which) {
    case SWAP_P_XP: {
      if (!n) return 0;
      double* d = nullptr;
      HIPCHK(c, hipMalloc(&d, 24 * n));
      TRY(pl_export_abs(c, d));
      HIPCHK(c, hipMemcpyAsync(host, d, 24 * n, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(d));
      return 0;
    }
    case SWAP_P_VP:
    case SWAP_P_VP1:
    case SWAP_P_FRAC: {
      int base = which == SWAP_P_VP ? 3 : (which == SWAP_P_VP1 ? 6 : 0);
      for (int a = 0; a < 3; a++)
        HIPCHK(c, hipMemcpyAsync((char*)host + a * rn, c->p[base + a], rn, hipMemcpyDeviceToHost,
                                 c->stream));
      break;
    }
    case SWAP_P_EP:
    case SWAP_P_VPLUS:
    case SWAP_P_VMINUS:
    case SWAP_P_WGHT:
    case SWAP_P_IJKP: {
      void* src = which == SWAP_P_EP ? c->Ep
                  : which == SWAP_P_VPLUS ? c->vpl
                  : which == SWAP_P_VMINUS ? c->vmi
                  : which == SWAP_P_WGHT ? c->wght : (void*)c->ijkp;
      if (!src) return set_err(c, SWAP_ESTATE, "particle array %d not computed yet", which);
      int nc = which == SWAP_P_WGHT ? 8 : 3;
      size_t es = which == SWAP_P_IJKP ? 4 : c->rsz;
      for (int a = 0; a < nc; a++)
        HIPCHK(c, hipMemcpyAsync((char*)host + a * es * n, (char*)src + a * es * c->cap, es * n,
                                 hipMemcpyDeviceToHost, c->stream));
      break;
    }
    case SWAP_P_CELL:
      HIPCHK(c, hipMemcpyAsync(host, c->cell, 4 * n, hipMemcpyDeviceToHost, c->stream));
      break;
    case SWAP_P_BETA:
    case SWAP_P_MRAT: {
      float* src = which == SWAP_P_BETA ? c->bp : c->mr;
      if (!src) {
        for (size_t q = 0; q < n; q++) ((float*)host)[q] = 1.0f;
        return 0;
      }
      HIPCHK(c, hipMemcpyAsync(host, src, 4 * n, hipMemcpyDeviceToHost, c->stream));
      break;
    }
    case SWAP_P_ID:
      if (!c->id) return set_err(c, SWAP_ESTATE, "ids are kept only with SWAP_TRACK_IDS");
      HIPCHK(c, hipMemcpyAsync(host, c->id, 4 * n, hipMemcpyDeviceToHost, c->stream));
      break;
    default:
      return set_err(c, SWAP_EINVAL, "unknown particle array %d", which);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swap_set_particle_array(swap_ctx* c, int which, const void* host) {
  if (!c || !host) return SWAP_EINVAL;
  TRY(compact(c));
  c->mgen++;
  size_t n = (size_t)c->n, rn = c->rsz * n;
  if (!n) return 0;
  switch (which) {
    case SWAP_P_XP: {
      double* d = nullptr;
      HIPCHK(c, hipMalloc(&d, 24 * n));
      HIPCHK(c, hipMemcpyAsync(d, host, 24 * n, hipMemcpyHostToDevice, c->stream));
      int e = pl_set_abs(c, d);
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(d));
      c->n_sorted = -1;
      return e;
    }
    case SWAP_P_VP:
    case SWAP_P_VP1:
    case SWAP_P_FRAC: {
      int base = which == SWAP_P_VP ? 3 : (which == SWAP_P_VP1 ? 6 : 0);
      for (int a = 0; a < 3; a++)
        HIPCHK(c, hipMemcpyAsync(c->p[base + a], (const char*)host + a * rn, rn,
                                 hipMemcpyHostToDevice, c->stream));
      if (which == SWAP_P_VP) c->vb_valid = false;
      break;
    }
    case SWAP_P_EP:
    case SWAP_P_VPLUS:
    case SWAP_P_VMINUS:
    case SWAP_P_WGHT:
    case SWAP_P_IJKP: {
      TRY(pl_ensure_scratch(c));
      void* dst = which == SWAP_P_EP ? c->Ep
                  : which == SWAP_P_VPLUS ? c->vpl
                  : which == SWAP_P_VMINUS ? c->vmi
                  : which == SWAP_P_WGHT ? c->wght : (void*)c->ijkp;
      int nc = which == SWAP_P_WGHT ? 8 : 3;
      size_t es = which == SWAP_P_IJKP ? 4 : c->rsz;
      for (int a = 0; a < nc; a++)
        HIPCHK(c, hipMemcpyAsync((char*)dst + a * es * c->cap, (const char*)host + a * es * n,
                                 es * n, hipMemcpyHostToDevice, c->stream));
      break;
    }
    case SWAP_P_CELL:
      HIPCHK(c, hipMemcpyAsync(c->cell, host, 4 * n, hipMemcpyHostToDevice, c->stream));
      c->n_sorted = -1;
      break;
    case SWAP_P_BETA:
    case SWAP_P_MRAT: {
      float* dst = which == SWAP_P_BETA ? c->bp : c->mr;
      if (!dst) return set_err(c, SWAP_ESTATE, "beta_p/mrat arrays need SWAP_SPECIES");
      HIPCHK(c, hipMemcpyAsync(dst, host, 4 * n, hipMemcpyHostToDevice, c->stream));
      break;
    }
    case SWAP_P_ID:
      if (!c->id) return set_err(c, SWAP_ESTATE, "ids are kept only with SWAP_TRACK_IDS");
      HIPCHK(c, hipMemcpyAsync(c->id, host, 4 * n, hipMemcpyHostToDevice, c->stream));
      {  // appended ions continue after the largest id given
        const uint32_t* h = (const uint32_t*)host;
        int64_t next = 0;
        for (size_t q = 0; q < n; q++) next = std::max<int64_t>(next, (int64_t)h[q] + 1);
        c->next_id = next;
      }
      break;
    default:
      return set_err(c, SWAP_EINVAL, "unknown particle array %d", which);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------- fields ----------------

int swap_set_field(swap_ctx* c, int f, const double* h) {
  if (!c || !h || f < 0 || f >= SWAP_F_COUNT) return SWAP_EINVAL;
  size_t n = (size_t)c->g.nn * ncomp_of(f);
  if (is_b(f)) {
    if (f == SWAP_F_B1P2) c->b1p2_alias = false;
    if (f == SWAP_F_B1 && c->b1p2_alias) {  // keep b1p2's current value
      TRY(fl_copy_b(c, c->b1, c->b1p2));
      c->b1p2_alias = false;
    }
    HIPCHK(c, hipMemcpyAsync(bptr(c, f), h, 8 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  return host_to_real(c, c->fld[f], n, h, nullptr);
}

int swap_get_field(swap_ctx* c, int f, double* h) {
  if (!c || !h || f < 0 || f >= SWAP_F_COUNT) return SWAP_EINVAL;
  size_t n = (size_t)c->g.nn * ncomp_of(f);
  if (is_b(f)) {
    HIPCHK(c, hipMemcpyAsync(h, bptr(c, f), 8 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  return real_to_host(c, c->fld[f], n, h, nullptr);
}

int swap_set_field_f32(swap_ctx* c, int f, const float* h) {
  if (!c || !h || f < 0 || f >= SWAP_F_COUNT) return SWAP_EINVAL;
  size_t n = (size_t)c->g.nn * ncomp_of(f);
  if (is_b(f)) {
    std::vector<double> t(h, h + n);
    return swap_set_field(c, f, t.data());
  }
  return host_to_real(c, c->fld[f], n, nullptr, h);
}

int swap_get_field_f32(swap_ctx* c, int f, float* h) {
  if (!c || !h || f < 0 || f >= SWAP_F_COUNT) return SWAP_EINVAL;
  size_t n = (size_t)c->g.nn * ncomp_of(f);
  if (is_b(f)) {
    std::vector<double> t(n);
    TRY(swap_get_field(c, f, t.data()));
    for (size_t q = 0; q < n; q++) h[q] = (float)t[q];
    return 0;
  }
  return real_to_host(c, c->fld[f], n, nullptr, h);
}

int swap_periodic(swap_ctx* c, int f) {
  if (!c || f < 0 || f >= SWAP_F_COUNT) return SWAP_EINVAL;
  if (is_b(f)) return fl_periodic(c, bptr(c, f), 3, 1);
  return fl_periodic(c, c->fld[f], ncomp_of(f), 0);
}

// ---------------- routine-level API ----------------

int swap_get_interp_weights(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(compact(c));
  return pl_weights(c);
}

int swap_update_np(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(compact(c));
  TRY(zero_moments(c, 0, 1));
  TRY(pl_deposit_simple(c, 0));
  return fl_finalize_np(c);
}

int swap_update_up(swap_ctx* c, int vsrc) {
  if (!c || vsrc < 0 || vsrc > 2) return SWAP_EINVAL;
  TRY(compact(c));
  TRY(zero_moments(c, 1, 4));
  TRY(pl_deposit_simple(c, vsrc == SWAP_V_CUR ? 1 : (vsrc == SWAP_V_EXTRAP ? 2 : 3)));
  return fl_finalize_up(c);
}

int swap_extrapol_up(swap_ctx* c) { return swap_update_up(c, SWAP_V_EXTRAP); }

int swap_get_Ep(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(compact(c));
  TRY(fl_build_src(c, c->fld[SWAP_F_AJ], c->fld[SWAP_F_UP], c->fld[SWAP_F_BTC],
                   c->fld[SWAP_F_AJC], c->S));
  return pl_get_Ep(c);
}

int swap_get_vplus_vminus(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(compact(c));
  if (!c->Ep) return set_err(c, SWAP_ESTATE, "get_vplus_vminus before get_Ep");
  return pl_vplus_vminus(c);
}

int swap_improve_up(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  if (!c->vpl) return set_err(c, SWAP_ESTATE, "improve_up before get_vplus_vminus");
  TRY(pl_improve_vp1(c));
  return swap_update_up(c, SWAP_V_VP1);
}

int swap_get_vp_final(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(compact(c));
  if (!c->vpl) return set_err(c, SWAP_ESTATE, "get_vp_final before get_vplus_vminus");
  c->mgen++;
  return pl_vp_final(c);
}

int swap_move_ion_half(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  if (c->g.ox)
    return set_err(c, SWAP_EINVAL, "move_ion_half wraps x; an open x moves only in swap_step");
  c->n_sorted = -1;  // boxes of the work items no longer describe the particles
  c->mgen++;
  return pl_move_half(c);
}

int swap_curlB(swap_ctx* c, int b_in, int aj_out) {
  if (!c || !is_b(b_in) || aj_out < 0 || aj_out >= SWAP_F_COUNT || is_b(aj_out) ||
      ncomp_of(aj_out) != 3)
    return SWAP_EINVAL;
  return fl_curlB(c, bptr(c, b_in), nullptr, c->fld[aj_out], nullptr);
}

int swap_curlE(swap_ctx* c, int e_in, int out) {
  if (!c || is_b(e_in) || is_b(out) || e_in < 0 || out < 0 || e_in >= SWAP_F_COUNT ||
      out >= SWAP_F_COUNT || ncomp_of(e_in) != 3 || ncomp_of(out) != 3 || e_in == out)
    return SWAP_EINVAL;
  return fl_curlE(c, c->fld[e_in], c->fld[out]);
}

int swap_crossf2(swap_ctx* c, int aa, int btc, int out) {
  if (!c || is_b(aa) || is_b(btc) || is_b(out) || aa < 0 || btc < 0 || out < 0 ||
      aa >= SWAP_F_COUNT || btc >= SWAP_F_COUNT || out >= SWAP_F_COUNT || out == aa ||
      out == btc)
    return SWAP_EINVAL;
  return fl_crossf2(c, c->fld[aa], c->fld[btc], c->fld[out]);
}

int swap_get_E(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  return get_E_impl(c);
}

int swap_predict_B(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(get_E_impl(c));
  TRY(fl_curlE_update(c, c->fld[SWAP_F_E], c->b12, 2.0 * c->dtsub, c->b1p2));
  c->b1p2_alias = false;
  return 0;
}

int swap_get_Ep1(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  return get_Ep1_impl(c);
}

int swap_correct_B(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(get_Ep1_impl(c));
  TRY(fl_curlE_update(c, c->fld[SWAP_F_E], c->b1, c->dtsub, c->b1p2));
  c->b1p2_alias = false;
  return 0;
}

int swap_f_update_tlev(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  return tlev_impl(c);
}

int swap_check_time_step(swap_ctx* c, int* ntf, double* dtsub) {
  if (!c) return SWAP_EINVAL;
  int n;
  double d;
  TRY(fl_check_time_step(c, &n, &d));
  c->ntf = n;
  c->dtsub = d;
  if (ntf) *ntf = n;
  if (dtsub) *dtsub = d;
  return 0;
}

int swap_edge_to_center(swap_ctx* c, int in, int out) {
  if (!c || in < 0 || out < 0 || in >= SWAP_F_COUNT || out >= SWAP_F_COUNT || is_b(out) ||
      ncomp_of(out) != 3 || ncomp_of(in) != 3 || in == out)
    return SWAP_EINVAL;
  if (is_b(in)) return fl_e2c_B(c, bptr(c, in), nullptr, c->fld[out]);
  return fl_e2c(c, c->fld[in], c->fld[out]);
}

int swap_edge_to_face(swap_ctx* c, int in, int out) {
  if (!c || in < 0 || out < 0 || in >= SWAP_F_COUNT || out >= SWAP_F_COUNT || is_b(out) ||
      ncomp_of(out) != 3 || ncomp_of(in) != 3 || in == out)
    return SWAP_EINVAL;
  if (is_b(in)) return fl_e2f(c, nullptr, bptr(c, in), c->fld[out]);
  return fl_e2f(c, c->fld[in], nullptr, c->fld[out]);
}

int swap_face_to_center(swap_ctx* c, int in, int out) {
  if (!c || in < 0 || out < 0 || in >= SWAP_F_COUNT || out >= SWAP_F_COUNT || is_b(out) ||
      is_b(in) || ncomp_of(out) != 3 || ncomp_of(in) != 3 || in == out)
    return SWAP_EINVAL;
  return fl_f2c(c, c->fld[in], c->fld[out]);
}

// ---------------- fused step ----------------

int swap_field_subcycles(swap_ctx* c, int ntf, double dtsub) {
  if (!c || ntf < 0 || !(dtsub > 0)) return SWAP_EINVAL;
  const bool fused = !(c->cfg.flags & SWAP_UNFUSED_FIELDS);
  if (fused) {
    TRY(fl_set_dtsub(c, dtsub, ntf));
    return subcycles_fused(c, ntf);
  }
  for (int q = 0; q < ntf; q++) TRY(subcycle_impl(c, dtsub));
  return 0;
}

int swap_sort(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  TRY(so_sort(c));
  if (!c->vb_valid) {
    TRY(sr_vbounds(c));
    c->vb_valid = true;
  }
  return 0;
}

int swap_prime(swap_ctx* c) {
  if (!c) return SWAP_EINVAL;
  if (c->cfg.sort_every > 0 && c->n_sorted != c->n) TRY(swap_sort(c));
  TRY(zero_moments(c, 0, 4));
  TRY(pl_pass(c, 1));
  TRY(fl_finalize_np(c));
  return fl_finalize_up(c);
}

int swap_step(swap_ctx* c, int nsteps) {
  if (!c || nsteps < 0) return SWAP_EINVAL;
  for (int s = 0; s < nsteps; s++) {
    // sort on schedule, or when particles were loaded (nothing sorted), or
    // when injected ions appended past the sorted range exceed 1/16 of all
    // (the passes take such a tail in small work items of consecutive ions,
    // whose windows the thermal spread of a few steps outgrows), or
    // when the last P3 found an item whose box outgrew its P1 window (its
    // particles would all deposit straight to HBM)
    bool due = c->cfg.sort_every > 0 &&
               (c->n_sorted < 0 || c->steps_since_sort >= c->cfg.sort_every ||
                (c->n - c->n_sorted) * 16 > c->n || (c->steps_since_sort > 0 && c->box_over > 0));
    if (due) {
      Stage t(c, SWAP_K_SORT);
      TRY(swap_sort(c));
    }
    // P1: np and extrapolated up at x^n (maind.f:385-399)
    TRY(zero_moments(c, 0, 4));
    {
      Stage t(c, SWAP_K_P1);
      TRY(pl_pass(c, 1));
    }
    {
      Stage t(c, SWAP_K_GLUE);
      TRY(fl_finalize_np(c));
      TRY(fl_finalize_up(c));
      TRY(dn_apply_step(c, true));  // learned denoiser, when enabled
      // node sources: aj = curlB(bt,np), btc, ajc - upc (maind.f:395-398, gutsp.f:183-184)
      TRY(fl_curlB(c, c->b1, nullptr, c->fld[SWAP_F_AJ], nullptr));
      TRY(fl_e2c_B(c, c->b1, nullptr, c->fld[SWAP_F_BTC]));
      TRY(fl_build_src(c, c->fld[SWAP_F_AJ], c->fld[SWAP_F_UP], c->fld[SWAP_F_BTC],
                       c->fld[SWAP_F_AJC], c->S));
      TRY(zero_moments(c, 1, 4));
    }
    // P2: improved up (maind.f:400-402)
    {
      Stage t(c, SWAP_K_P2);
      TRY(pl_pass(c, 2));
    }
    {
      Stage t(c, SWAP_K_GLUE);
      TRY(fl_finalize_up(c));
      TRY(dn_apply_step(c, false));
      TRY(fl_update_src(c, c->fld[SWAP_F_UP], c->fld[SWAP_F_AJC], c->S));
      TRY(zero_moments(c, 0, 4));
      TRY(pl_box_reset(c));
    }
    // P3: final velocity, both half moves, moments at x^{n+1/2} (maind.f:404-413, 479)
    {
      Stage t(c, SWAP_K_P3);
      TRY(pl_pass(c, 3));
      if (c->g.ox) c->dead_pending = true;  // P3 marks the ions that left
      TRY(sr_inflow_step(c));  // an open x: the ions entering during this step
    }
    {
      Stage t(c, SWAP_K_GLUE);
      TRY(fl_finalize_np(c));
      TRY(fl_finalize_up(c));
      TRY(dn_apply_step(c, true));
    }
    // P3 reduced the cell boxes of x^{n+1} and the bounds of v^{n+1/2}
    std::swap(c->vb, c->vb_next);
    c->vb_valid = c->n > 0;  // the inflow ions' bounds were folded in
    if (c->n_sorted >= 0) {
      std::swap(c->boxes, c->boxes_next);
      c->steps_since_sort++;
    }
    // field subcycles (maind.f:420-476)
    if (c->cfg.flags & SWAP_UNFUSED_FIELDS) {  // per-routine kernels with host dtsub
      int ntf;
      double dtsub;
      TRY(swap_check_time_step(c, &ntf, &dtsub));
      Stage t(c, SWAP_K_FIELD);
      TRY(swap_field_subcycles(c, ntf, dtsub));
    } else {
      // check_time_step on the device: the subcycle kernels read dtsub there,
      // so the usual ntsub subcycles are enqueued at once; ntf (and the sort
      // trigger of the next step) come back through pinned memory, long
      // before those subcycles finish, and a step that halved dtsub enqueues
      // the rest
      TRY(fl_time_step_dev(c));
      HIPCHK(c, hipMemcpyAsync(&c->hstat->dtsub, c->ts, 2 * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(&c->hstat->box_over, c->dcnt + 4, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipEventRecord(c->ev_stat, c->stream));
      Stage t(c, SWAP_K_FIELD);
      const int n0 = c->cfg.ntsub;
      TRY(subcycles_fused(c, n0));
      HIPCHK(c, hipEventSynchronize(c->ev_stat));
      const int ntf = (int)c->hstat->ntf;
      c->box_over = c->hstat->box_over;
      for (int q = n0; q < ntf; q++) TRY(subcycle_fused(c, q == ntf - 1));
      c->ntf = ntf;
      c->dtsub = c->hstat->dtsub;
    }
    c->step_count++;
    c->step_gen = c->mgen;  // macc: the final deposit of this step
  }
  return 0;
}

int swap_ionize_sw(swap_ctx* c, int step, double ndot, double vsw, double beta_pu,
                   int64_t* injected) {
  if (!c || !injected) return SWAP_EINVAL;
  c->mgen++;
  return sr_ionize_sw(c, step, ndot, vsw, beta_pu, injected);
}

int swap_set_inflow(swap_ctx* c, double ppc, double vsw, double vth, double beta_p) {
  if (!c) return SWAP_EINVAL;
  return sr_set_inflow(c, ppc, vsw, vth, beta_p);
}

int swap_ionize_cloud(swap_ctx* c, int step, const swap_cloud* cloud, int64_t* injected,
                      int64_t* chex) {
  if (!c || !cloud || !injected || !chex) return SWAP_EINVAL;
  c->mgen++;
  return sr_ionize_cloud(c, step, cloud, injected, chex);
}

int swap_diagnostics(swap_ctx* c, swap_diag* d) {
  if (!c || !d) return SWAP_EINVAL;
  TRY(compact(c));
  return sr_diagnostics(c, d);
}

int swap_energy_diag(swap_ctx* c, double mion, const double b0[3], swap_energy* d) {
  if (!c || !b0 || !d || !(mion > 0)) return SWAP_EINVAL;
  TRY(compact(c));
  return sr_energy_diag(c, mion, b0, d);
}

int swap_set_input_energy(swap_ctx* c, double mion, double input_E) {
  if (!c || !(mion > 0) || !std::isfinite(input_E)) return SWAP_EINVAL;
  double k = input_E / mion;
  HIPCHK(c, hipMemcpyAsync(c->dred + DRED_INPUT, &k, sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------- run state (checkpoint / resume) ----------------

namespace {
constexpr char RS_MAGIC[8] = {'S', 'W', 'A', 'P', 'R', 'S', '0', '1'};
struct RunStateHead {
  char magic[8];
  int64_t max_items, vb_ints, ntiles;  // shape of the arrays that follow
  int64_t n, step_count, steps_since_sort, n_sorted, next_id, injected, escaped, sorts;
  uint64_t box_over;
  double in_ppc, in_vsw, in_vth, in_bp, input_k;
  int32_t vb_valid, nitems;
  uint64_t counters[8];  // dcnt: window misses, escapes, scratch, chex, box overflows, ...
};
size_t run_state_bytes(const swap_ctx* c) {
  return sizeof(RunStateHead) + sizeof(Item) * (size_t)c->max_items +
         24 * (size_t)c->max_items + sizeof(int) * (size_t)vb_ints(c);
}
}  // namespace

int swap_run_state_size(swap_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return SWAP_EINVAL;
  *bytes = (int64_t)run_state_bytes(c);
  return 0;
}

int swap_save_run_state(swap_ctx* c, void* host) {
  if (!c || !host) return SWAP_EINVAL;
  TRY(compact(c));
  RunStateHead h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, RS_MAGIC, 8);
  h.max_items = c->max_items;
  h.vb_ints = vb_ints(c);
  h.ntiles = c->ntiles;
  h.n = c->n;
  h.step_count = c->step_count;
  h.steps_since_sort = c->steps_since_sort;
  h.n_sorted = c->n_sorted;
  h.next_id = c->next_id;
  h.injected = c->injected;
  h.escaped = c->escaped;
  h.sorts = c->sorts;
  h.box_over = c->box_over;
  h.in_ppc = c->in_ppc;
  h.in_vsw = c->in_vsw;
  h.in_vth = c->in_vth;
  h.in_bp = c->in_bp;
  h.vb_valid = c->vb_valid ? 1 : 0;
  char* b = (char*)host + sizeof h;
  HIPCHK(c, hipMemcpyAsync(&h.input_k, c->dred + DRED_INPUT, sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&h.nitems, c->d_nitems, sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(h.counters, c->dcnt, sizeof h.counters, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(b, c->items, sizeof(Item) * (size_t)c->max_items,
                           hipMemcpyDeviceToHost, c->stream));
  b += sizeof(Item) * (size_t)c->max_items;
  HIPCHK(c, hipMemcpyAsync(b, c->boxes, 24 * (size_t)c->max_items, hipMemcpyDeviceToHost,
                           c->stream));
  b += 24 * (size_t)c->max_items;
  HIPCHK(c, hipMemcpyAsync(b, c->vb, sizeof(int) * (size_t)vb_ints(c), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(host, &h, sizeof h);
  return 0;
}

int swap_restore_run_state(swap_ctx* c, const void* host) {
  if (!c || !host) return SWAP_EINVAL;
  RunStateHead h;
  memcpy(&h, host, sizeof h);
  if (memcmp(h.magic, RS_MAGIC, 8) != 0)
    return set_err(c, SWAP_EINVAL, "restore_run_state: not a run state");
  if (h.max_items != c->max_items || h.vb_ints != vb_ints(c) || h.ntiles != c->ntiles)
    return set_err(c, SWAP_EINVAL, "restore_run_state: saved from a context of another shape");
  if (h.n != c->n)
    return set_err(c, SWAP_ESTATE,
                   "restore_run_state: saved with %lld particles, %lld loaded (load the "
                   "particles first)", (long long)h.n, (long long)c->n);
  if (h.n_sorted > c->n || h.n_sorted < -1 || h.nitems < 0 || h.nitems > c->max_items)
    return set_err(c, SWAP_EINVAL, "restore_run_state: inconsistent sort state");
  if (h.in_ppc > 0 || c->in_ppc > 0)
    TRY(sr_set_inflow(c, h.in_ppc, h.in_vsw, h.in_vth, h.in_bp));
  const char* b = (const char*)host + sizeof h;
  HIPCHK(c, hipMemcpyAsync(c->items, b, sizeof(Item) * (size_t)c->max_items,
                           hipMemcpyHostToDevice, c->stream));
  b += sizeof(Item) * (size_t)c->max_items;
  HIPCHK(c, hipMemcpyAsync(c->boxes, b, 24 * (size_t)c->max_items, hipMemcpyHostToDevice,
                           c->stream));
  b += 24 * (size_t)c->max_items;
  HIPCHK(c, hipMemcpyAsync(c->vb, b, sizeof(int) * (size_t)vb_ints(c), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_nitems, &h.nitems, sizeof(int), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->dcnt, h.counters, sizeof h.counters, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->dred + DRED_INPUT, &h.input_k, sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->step_count = h.step_count;
  c->steps_since_sort = (int)h.steps_since_sort;
  c->n_sorted = h.n_sorted;
  c->next_id = h.next_id;
  c->injected = h.injected;
  c->escaped = h.escaped;
  c->sorts = h.sorts;
  c->box_over = h.box_over;
  c->vb_valid = h.vb_valid != 0;
  c->mgen++;
  return 0;
}

int swap_profile(swap_ctx* c, int enable) {
  if (!c) return SWAP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  prof_clear(c);
  c->prof_on = enable != 0;
  return 0;
}

int swap_profile_read(swap_ctx* c, double* ms, int64_t* count, int n) {
  if (!c || !ms || !count || n < 0) return SWAP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < n && k < SWAP_K_COUNT; k++) {
    double t = 0;
    const auto& ev = c->prof_ev[k];
    for (size_t q = 0; q + 1 < ev.size(); q += 2) {
      float e = 0;
      HIPCHK(c, hipEventElapsedTime(&e, ev[q], ev[q + 1]));
      t += e;
    }
    ms[k] = t;
    count[k] = (int64_t)(ev.size() / 2);
  }
  return 0;
}

}  // extern "C"
