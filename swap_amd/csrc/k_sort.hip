// k_sort.hip -- stable counting sort of particles by (tile, cell in tile).
//
// Build-only component (the reference keeps particles in load order).  The
// key of a particle is tile * TILE_CELLS + the cell's linear index inside its
// tile, so a work item (a run of one tile's particles) touches a compact
// block of nodes.
//
// The sort works on runs: a run is a maximal group of consecutive particles
// inside one wavefront that share a key (input that was sorted a few steps
// ago is made of long runs).
//   A: count particles and runs per key (one atomic per run);
//   scans of both counts;
//   B: write each run (head index, length) into its key's run list;
//   C: per key, order its runs by head index and turn lengths into
//      destination offsets;
//   D: every particle finds its run's offset and is written once, with all
//      its arrays, to the destination.
// The result is exactly the stable order (runs of a key ordered by their
// first index, indices inside a run consecutive), independent of the order
// in which atomics were served.
#include "swap_internal.h"

namespace {

__device__ __forceinline__ uint32_t tile_key(uint32_t c, int ntx, int nty) {
  int i = cell_i(c), j = cell_j(c), k = cell_k(c);
  uint32_t tile = (uint32_t)(i / TILE_X) + (uint32_t)ntx * ((uint32_t)(j / TILE_Y) +
                                                            (uint32_t)nty * (uint32_t)(k / TILE_Z));
  return tile * TILE_CELLS + (uint32_t)(i % TILE_X) +
         TILE_X * ((uint32_t)(j % TILE_Y) + TILE_Y * (uint32_t)(k % TILE_Z));
}

struct Run {
  int head;  // lane of this lane's run head
  int len;   // run length (meaningful on the head lane)
};
__device__ __forceinline__ Run wave_runs(uint32_t key) {
  const int lane = __lane_id();
  uint32_t prev = __shfl_up(key, 1);
  bool head = lane == 0 || prev != key;
  unsigned long long hm = __ballot(head);
  unsigned long long le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
  Run r;
  r.head = 63 - __clzll(hm & le);
  unsigned long long above = hm & ~le;
  int nh = above ? __ffsll((long long)above) - 1 : 64;
  r.len = nh - lane;
  return r;
}

__device__ __forceinline__ uint32_t key_at(int64_t p, int64_t n, const uint32_t* cell, int ntx,
                                           int nty) {
  return p < n ? tile_key(cell[p], ntx, nty) : 0xFFFFFFFFu;
}

__global__ void k_sortA(int64_t n, const uint32_t* cell, int ntx, int nty, uint32_t* cnt,
                        uint32_t* nrun) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  uint32_t k = key_at(p, n, cell, ntx, nty);
  Run r = wave_runs(k);
  if (p < n && r.head == __lane_id()) {
    atomicAdd(cnt + k, (uint32_t)r.len);
    atomicAdd(nrun + k, 1u);
  }
}

__global__ void k_sortB(int64_t n, const uint32_t* cell, int ntx, int nty, const uint32_t* rstart,
                        uint32_t* rcur, uint2* runs) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  uint32_t k = key_at(p, n, cell, ntx, nty);
  Run r = wave_runs(k);
  if (p < n && r.head == __lane_id()) {
    uint32_t slot = rstart[k] + atomicAdd(rcur + k, 1u);
    runs[slot] = make_uint2((uint32_t)p, (uint32_t)r.len);
  }
}

// one thread per run: its destination is the key's start plus the lengths of
// the key's runs with a smaller head index; stored at the head's index
__global__ void k_sortC(const uint32_t* cell, int ntx, int nty, const uint32_t* kstart,
                        const uint32_t* rstart, int64_t nkeys, const uint2* runs,
                        uint32_t* rbase) {
  const uint32_t nruns = rstart[nkeys];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nruns;
       q += (int64_t)gridDim.x * blockDim.x) {
    uint2 me = runs[q];
    uint32_t k = tile_key(cell[me.x], ntx, nty);
    uint32_t base = kstart[k];
    for (uint32_t x = rstart[k]; x < rstart[k + 1]; x++) {
      uint2 o = runs[x];
      if (o.x < me.x) base += o.y;
    }
    rbase[me.x] = base;
  }
}

struct Move {
  const uint32_t* cell;
  uint32_t* cell2;
  const void* p[9];
  void* p2[9];
  const float *bp, *mr;
  float *bp2, *mr2;
  const uint32_t* id;
  uint32_t* id2;
};

template <typename R>
__global__ void k_sortD(int64_t n, int ntx, int nty, const uint32_t* rbase, Move Mv) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  uint32_t c = p < n ? Mv.cell[p] : 0u;
  uint32_t k = p < n ? tile_key(c, ntx, nty) : 0xFFFFFFFFu;
  Run r = wave_runs(k);
  uint32_t base = 0;
  if (p < n && r.head == __lane_id()) base = rbase[p];
  base = __shfl(base, r.head);
  if (p >= n) return;
  uint32_t d = base + (uint32_t)(__lane_id() - r.head);
  Mv.cell2[d] = c;
#pragma unroll
  for (int a = 0; a < 9; a++) ((R*)Mv.p2[a])[d] = ((const R*)Mv.p[a])[p];
  if (Mv.bp) Mv.bp2[d] = Mv.bp[p];
  if (Mv.mr) Mv.mr2[d] = Mv.mr[p];
  if (Mv.id) Mv.id2[d] = Mv.id[p];
}

// ---- exclusive scan of uint32 counts (3 phases) ----
constexpr int SCAN_T = 1024, SCAN_E = 4;
__global__ void k_scan_blocks(int64_t n, const uint32_t* in, uint32_t* out, uint32_t* bsum) {
  __shared__ uint32_t s[SCAN_T];
  int64_t base = blockIdx.x * (int64_t)(SCAN_T * SCAN_E) + (int64_t)threadIdx.x * SCAN_E;
  uint32_t v[SCAN_E], t = 0;
  for (int e = 0; e < SCAN_E; e++) {
    v[e] = base + e < n ? in[base + e] : 0;
    t += v[e];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    uint32_t x = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - t;
  for (int e = 0; e < SCAN_E; e++) {
    if (base + e < n) out[base + e] = run;
    run += v[e];
  }
  if (threadIdx.x == SCAN_T - 1) bsum[blockIdx.x] = s[threadIdx.x];
}
// single block: exclusive scan of the block sums in place, total -> *total
__global__ void k_scan_top(int64_t nb, uint32_t* bsum, uint32_t* total) {
  __shared__ uint32_t s[SCAN_T];
  int64_t per = (nb + SCAN_T - 1) / SCAN_T;
  int64_t b0 = threadIdx.x * per;
  uint32_t t = 0;
  for (int64_t q = b0; q < b0 + per && q < nb; q++) t += bsum[q];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    uint32_t x = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - t;
  for (int64_t q = b0; q < b0 + per && q < nb; q++) {
    uint32_t v = bsum[q];
    bsum[q] = run;
    run += v;
  }
  if (threadIdx.x == SCAN_T - 1) *total = s[threadIdx.x];
}
__global__ void k_scan_add(int64_t n, uint32_t* out, const uint32_t* bsum) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  out[q] += bsum[q / (SCAN_T * SCAN_E)];
}

// work items: each tile's run is cut into pieces of at most ITEM_MAX; the
// box of an item is the tile's x-y extent and the z layers of its first and
// last particle (sorted cells are in tile-local linear order)
__global__ void k_items_count(int64_t ntiles, const uint32_t* start, uint32_t* nit) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t a = start[t * TILE_CELLS], b = start[(t + 1) * TILE_CELLS];
  nit[t] = (b - a + ITEM_MAX - 1) / ITEM_MAX;
}
__global__ void k_items_fill(int64_t ntiles, const uint32_t* start, const uint32_t* nit,
                             const uint32_t* ioff, const uint32_t* cell_sorted, int ntx, int nty,
                             int ncx, int ncy, Item* items, int* boxes) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t a = start[t * TILE_CELLS], b = start[(t + 1) * TILE_CELLS];
  int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty);
  for (uint32_t q = 0; q < nit[t]; q++) {
    uint32_t s = a + q * ITEM_MAX, e = min(b, s + ITEM_MAX);
    uint32_t w = ioff[t] + q;
    items[w].start = (int32_t)s;
    items[w].end = (int32_t)e;
    items[w].tile = (int32_t)t;
    items[w].pad = 0;
    int* bx = boxes + 6 * w;
    bx[0] = tx * TILE_X;
    bx[1] = ty * TILE_Y;
    bx[2] = cell_k(cell_sorted[s]);
    bx[3] = min((tx + 1) * TILE_X, ncx) - 1;
    bx[4] = min((ty + 1) * TILE_Y, ncy) - 1;
    bx[5] = cell_k(cell_sorted[e - 1]);
  }
}

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

// exclusive scan of n counts into out (n+1 entries: out[n] = total)
int scan(swap_ctx* c, int64_t n, const uint32_t* in, uint32_t* out) {
  int64_t nb = (n + SCAN_T * SCAN_E - 1) / (SCAN_T * SCAN_E);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SCAN_T), 0, c->stream, n, in, out,
                     c->sblk);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, c->stream, nb, c->sblk, out + n);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, out, c->sblk);
  return post(c);
}

}  // namespace

int so_sort(swap_ctx* c) {
  const int64_t n = c->n, nk = c->nkeys;
  const int ntx = c->ntile[0], nty = c->ntile[1];
  if (n > 0) {
    HIPCHK(c, hipMemsetAsync(c->scnt, 0, sizeof(uint32_t) * nk, c->stream));
    HIPCHK(c, hipMemsetAsync(c->snrun, 0, sizeof(uint32_t) * nk, c->stream));
    hipLaunchKernelGGL(k_sortA, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, c->cell, ntx, nty,
                       c->scnt, c->snrun);
    if (int e = post(c)) return e;
    if (int e = scan(c, nk, c->scnt, c->sstart)) return e;
    if (int e = scan(c, nk, c->snrun, c->srstart)) return e;
    HIPCHK(c, hipMemsetAsync(c->scur, 0, sizeof(uint32_t) * nk, c->stream));
    hipLaunchKernelGGL(k_sortB, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, c->cell, ntx, nty,
                       c->srstart, c->scur, c->sruns);
    hipLaunchKernelGGL(k_sortC, dim3(8192), dim3(256), 0, c->stream, c->cell, ntx, nty,
                       c->sstart, c->srstart, nk, c->sruns, c->srbase);
    Move Mv;
    Mv.cell = c->cell; Mv.cell2 = c->cell2;
    for (int a = 0; a < 9; a++) { Mv.p[a] = c->p[a]; Mv.p2[a] = c->p2[a]; }
    Mv.bp = c->bp; Mv.bp2 = c->bp2; Mv.mr = c->mr; Mv.mr2 = c->mr2;
    Mv.id = c->id; Mv.id2 = c->id2;
    if (c->prec == 64)
      hipLaunchKernelGGL(k_sortD<double>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, ntx,
                         nty, c->srbase, Mv);
    else
      hipLaunchKernelGGL(k_sortD<float>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, ntx,
                         nty, c->srbase, Mv);
    if (int e = post(c)) return e;
    std::swap(c->cell, c->cell2);
    for (int a = 0; a < 9; a++) std::swap(c->p[a], c->p2[a]);
    std::swap(c->bp, c->bp2);
    std::swap(c->mr, c->mr2);
    std::swap(c->id, c->id2);
  } else {
    HIPCHK(c, hipMemsetAsync(c->sstart, 0, sizeof(uint32_t) * (nk + 1), c->stream));
  }
  // work items and their cell boxes; the item count stays on the device
  uint32_t* nit = c->scur;  // reuse: ntiles + 1 entries each
  uint32_t* ioff = c->scur + c->ntiles + 1;
  hipLaunchKernelGGL(k_items_count, dim3(nblk(c->ntiles, 256)), dim3(256), 0, c->stream,
                     c->ntiles, c->sstart, nit);
  if (int e = scan(c, c->ntiles, nit, ioff)) return e;
  hipLaunchKernelGGL(k_items_fill, dim3(nblk(c->ntiles, 256)), dim3(256), 0, c->stream, c->ntiles,
                     c->sstart, nit, ioff, c->cell, ntx, nty, c->cfg.ncx, c->cfg.ncy, c->items,
                     c->boxes);
  HIPCHK(c, hipMemcpyAsync(c->d_nitems, ioff + c->ntiles, sizeof(int), hipMemcpyDeviceToDevice,
                           c->stream));
  if (int e = post(c)) return e;
  c->n_sorted = n;
  c->steps_since_sort = 0;
  return 0;
}
