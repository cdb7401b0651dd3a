// k_sort.hip -- counting sort of particles by (tile, cell in tile).
//
// Build-only component (the reference keeps particles in load order).  The
// key of a particle is tile * TILE_CELLS + the cell's linear index inside
// its tile, so a work item (a run of one tile's particles) touches a
// compact block of nodes.  The result is the stable sort: the histogram and
// scatter use run-aggregated atomics (one atomic per run of equal keys in a
// wavefront), and each key's segment is then restored to input order, so
// the permutation is bit-identical to a sequential stable counting sort.
#include "swap_internal.h"

namespace {

__device__ __forceinline__ uint32_t tile_key(uint32_t c, int ntx, int nty) {
  int i = cell_i(c), j = cell_j(c), k = cell_k(c);
  uint32_t tile = (uint32_t)(i / TILE_X) + (uint32_t)ntx * ((uint32_t)(j / TILE_Y) +
                                                            (uint32_t)nty * (uint32_t)(k / TILE_Z));
  return tile * TILE_CELLS + (uint32_t)(i % TILE_X) +
         TILE_X * ((uint32_t)(j % TILE_Y) + TILE_Y * (uint32_t)(k % TILE_Z));
}

// run structure of equal keys inside one wavefront
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

__global__ void k_keys_hist(int64_t n, const uint32_t* cell, int ntx, int nty, uint32_t* key,
                            uint32_t* cnt) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool valid = p < n;
  uint32_t k = 0xFFFFFFFFu;
  if (valid) {
    k = tile_key(cell[p], ntx, nty);
    key[p] = k;
  }
  Run r = wave_runs(k);
  if (valid && r.head == __lane_id()) atomicAdd(cnt + k, (uint32_t)r.len);
}

__global__ void k_scatter(int64_t n, const uint32_t* key, uint32_t* cur, uint32_t* perm) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool valid = p < n;
  uint32_t k = valid ? key[p] : 0xFFFFFFFFu;
  Run r = wave_runs(k);
  uint32_t base = 0;
  if (valid && r.head == __lane_id()) base = atomicAdd(cur + k, (uint32_t)r.len);
  base = __shfl(base, r.head);
  if (valid) perm[base + (__lane_id() - r.head)] = (uint32_t)p;
}

// restore input order inside each key segment (insertion sort; a segment is
// a concatenation of a few increasing runs)
__global__ void k_fix(int64_t nkeys, const uint32_t* start, uint32_t* perm) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nkeys) return;
  uint32_t a = start[q], b = start[q + 1];
  for (uint32_t x = a + 1; x < b; x++) {
    uint32_t v = perm[x];
    uint32_t y = x;
    while (y > a && perm[y - 1] > v) {
      perm[y] = perm[y - 1];
      y--;
    }
    perm[y] = v;
  }
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

struct Perm {
  uint32_t *cell, *cell2;
  void *p[9], *p2[9];
  float *bp, *bp2, *mr, *mr2;
  uint32_t *id, *id2;
};
template <typename R>
__global__ void k_permute(int64_t n, const uint32_t* perm, Perm P) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t s = perm[r];
  P.cell2[r] = P.cell[s];
#pragma unroll
  for (int a = 0; a < 9; a++) ((R*)P.p2[a])[r] = ((const R*)P.p[a])[s];
  if (P.bp) P.bp2[r] = P.bp[s];
  if (P.mr) P.mr2[r] = P.mr[s];
  if (P.id) P.id2[r] = P.id[s];
}

// work items: each tile's run is cut into pieces of at most ITEM_MAX
__global__ void k_items_count(int64_t ntiles, const uint32_t* start, uint32_t* nit) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t a = start[t * TILE_CELLS], b = start[(t + 1) * TILE_CELLS];
  nit[t] = (b - a + ITEM_MAX - 1) / ITEM_MAX;
}
__global__ void k_items_fill(int64_t ntiles, const uint32_t* start, const uint32_t* nit,
                             const uint32_t* ioff, const uint32_t* key, const uint32_t* perm,
                             int ntx, int nty, int ncx,
                             int ncy, int ncz, Item* items, int* boxes) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t a = start[t * TILE_CELLS], b = start[(t + 1) * TILE_CELLS];
  int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), tz = (int)(t / ((int64_t)ntx * nty));
  for (uint32_t q = 0; q < nit[t]; q++) {
    uint32_t s = a + q * ITEM_MAX, e = min(b, s + ITEM_MAX);
    uint32_t w = ioff[t] + q;
    items[w].start = (int32_t)s;
    items[w].end = (int32_t)e;
    items[w].tile = (int32_t)t;
    items[w].pad = 0;
    // cells of the run: the tile's x-y extent, z layers of its first/last key
    int lz0 = (int)((key[perm[s]] % TILE_CELLS) / (TILE_X * TILE_Y));
    int lz1 = (int)((key[perm[e - 1]] % TILE_CELLS) / (TILE_X * TILE_Y));
    int* bx = boxes + 6 * w;
    bx[0] = tx * TILE_X;
    bx[1] = ty * TILE_Y;
    bx[2] = tz * TILE_Z + lz0;
    bx[3] = min((tx + 1) * TILE_X, ncx) - 1;
    bx[4] = min((ty + 1) * TILE_Y, ncy) - 1;
    bx[5] = tz * TILE_Z + lz1;
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
  int64_t n = c->n;
  HIPCHK(c, hipMemsetAsync(c->scnt, 0, sizeof(uint32_t) * c->nkeys, c->stream));
  if (n > 0)
    hipLaunchKernelGGL(k_keys_hist, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, c->cell,
                       c->ntile[0], c->ntile[1], c->skey, c->scnt);
  if (int e = post(c)) return e;
  if (int e = scan(c, c->nkeys, c->scnt, c->sstart)) return e;
  HIPCHK(c, hipMemcpyAsync(c->scur, c->sstart, sizeof(uint32_t) * c->nkeys,
                           hipMemcpyDeviceToDevice, c->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, c->skey,
                       c->scur, c->sperm);
    hipLaunchKernelGGL(k_fix, dim3(nblk(c->nkeys, 256)), dim3(256), 0, c->stream, c->nkeys,
                       c->sstart, c->sperm);
    Perm P;
    P.cell = c->cell; P.cell2 = c->cell2;
    for (int a = 0; a < 9; a++) { P.p[a] = c->p[a]; P.p2[a] = c->p2[a]; }
    P.bp = c->bp; P.bp2 = c->bp2; P.mr = c->mr; P.mr2 = c->mr2;
    P.id = c->id; P.id2 = c->id2;
    if (c->prec == 64)
      hipLaunchKernelGGL(k_permute<double>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n,
                         c->sperm, P);
    else
      hipLaunchKernelGGL(k_permute<float>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n,
                         c->sperm, P);
    if (int e = post(c)) return e;
    std::swap(c->cell, c->cell2);
    for (int a = 0; a < 9; a++) std::swap(c->p[a], c->p2[a]);
    std::swap(c->bp, c->bp2);
    std::swap(c->mr, c->mr2);
    std::swap(c->id, c->id2);
  }
  // work items and their cell boxes; the item count stays on the device
  uint32_t* nit = c->scur;                // reuse: ntiles entries
  uint32_t* ioff = c->scur + c->ntiles + 1;
  hipLaunchKernelGGL(k_items_count, dim3(nblk(c->ntiles, 256)), dim3(256), 0, c->stream,
                     c->ntiles, c->sstart, nit);
  if (int e = scan(c, c->ntiles, nit, ioff)) return e;
  hipLaunchKernelGGL(k_items_fill, dim3(nblk(c->ntiles, 256)), dim3(256), 0, c->stream, c->ntiles,
                     c->sstart, nit, ioff, c->skey, c->sperm, c->ntile[0], c->ntile[1], c->cfg.ncx,
                     c->cfg.ncy, c->cfg.ncz, c->items, c->boxes);
  HIPCHK(c, hipMemcpyAsync(c->d_nitems, ioff + c->ntiles, sizeof(int), hipMemcpyDeviceToDevice,
                           c->stream));
  if (int e = post(c)) return e;
  c->n_sorted = n;
  c->steps_since_sort = 0;
  return 0;
}
