// k_sort.hip -- stable sort of particles by tile, cells in order within
// each source chunk.
//
// Build-only component (the reference keeps particles in load order).  The
// tiles are blocks of TILE_X x TILE_Y x TILE_Z cells; a work item (a run of
// one tile's particles) touches a compact block of nodes.
//
// The sort is two-level.  The particles are cut into chunks of SCH
// consecutive indices; a chunk's fine keys (tile * TILE_CELLS + the cell's
// linear index inside its tile) are first sorted stably inside LDS (a 4-bit
// LSD radix sort), which turns the chunk into segments, one per distinct
// tile, each in cell order.  Globally only segments are counted:
//   A: per chunk, LDS sort; record each segment (tile, start, length), count
//      particles and segments per tile (two atomics per segment) and store
//      each sorted position's source index and segment;
//   scans of both counts;
//   B: write each segment (chunk position, length) into its tile's list;
//   C: per tile, order its segments by chunk and turn lengths into
//      destination offsets;
//   D: per chunk, walk the sorted positions: every particle is written
//      once, with all its arrays, to its segment's offset + its rank, so
//      the writes of a segment are contiguous.
// The result is the stable order of the key (tile, source chunk, cell):
// a chunk holds at most one segment per tile, so ordering a tile's segments
// by chunk gives it independent of the order in which atomics were served.
// A tile's particles are thus a few runs in cell order (one per chunk that
// held some), which keeps neighbours in a cell together for the deposit,
// while a segment is a chunk's whole share of a tile: the particles drift
// with the flow and several cells between sorts, so segments by cell would
// hold 10-20 particles (partial cache lines at every segment's ends, written
// by different workgroups) where segments by tile hold hundreds.
#include "swap_internal.h"

namespace {

constexpr int ST = 512, SE = 8, SCH = ST * SE;  // chunk: 4096 particles
constexpr int SCNT = 16 * ST;                   // radix counters, bin-major
__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }

// exclusive scan of one value per thread over the workgroup; *total gets the
// sum.  Ends with a barrier, so wsum may be reused right away.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum,
                                                     uint32_t* total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (t == 0) {
    uint32_t s = 0;
    for (int q = 0; q < ST / 64; q++) {
      uint32_t u = wsum[q];
      wsum[q] = s;
      s += u;
    }
    wsum[ST / 64] = s;
  }
  __syncthreads();
  uint32_t r = x - v + wsum[w];
  *total = wsum[ST / 64];
  __syncthreads();
  return r;
}

// A: one workgroup per chunk.  The radix sort runs on a chunk-relative key
// that keeps the order of (tile, cell in tile): the tile's x, y and z
// offsets from the chunk's smallest, each in as many bits as the chunk's
// range of that coordinate needs, above the cell's 8-bit index in its tile.
// A chunk holds particles of a tile or two in each direction, so the key
// has about 14 bits (4 passes) where the absolute key would need up to 22 (a
// chunk at a periodic seam spans the grid and takes more).  Ions that left
// through an open x sort behind all live ones (the relative tile one past
// the largest), padding positions last.
__device__ __forceinline__ int nbits_of(uint32_t v) { return v ? 32 - __clz((int)v) : 0; }
__global__ __launch_bounds__(ST) void k_sortA(int64_t n, const uint32_t* __restrict__ cell,
                                              int ntx, int nty, uint32_t dead_tile,
                                              uint32_t* cnt,
                                              uint32_t* nrun, uint32_t* segkey,
                                              uint32_t* seginfo, uint32_t* nsegc,
                                              uint32_t* perm) {
  __shared__ uint32_t skey[SCH];
  __shared__ uint16_t sidx[SCH];
  __shared__ uint32_t scan[SCNT + SCNT / 16];
  __shared__ uint32_t wsum[ST / 64 + 1];
  __shared__ int red[6][ST / 64];
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * SCH;
  const int m = (int)min((int64_t)SCH, n - c0);
  uint32_t cs[SE], key[SE], idx[SE];
  if (t * SE + SE <= m) {
    const uint4* src = (const uint4*)(cell + c0) + 2 * t;  // c0 is a multiple of 4096
    uint4 a = src[0], b = src[1];
    cs[0] = a.x; cs[1] = a.y; cs[2] = a.z; cs[3] = a.w;
    cs[4] = b.x; cs[5] = b.y; cs[6] = b.z; cs[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < SE; e++) {
      int l = t * SE + e;
      cs[e] = l < m ? cell[c0 + l] : DEAD_CELL;
    }
  }
#pragma unroll
  for (int e = 0; e < SE; e++) idx[e] = (uint32_t)(t * SE + e);

  // the chunk's range of tile coordinates (live particles only)
  int mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, mx[3] = {-1, -1, -1};
#pragma unroll
  for (int e = 0; e < SE; e++)
    if (cs[e] != DEAD_CELL) {
      const int tc[3] = {cell_i(cs[e]) / TILE_X, cell_j(cs[e]) / TILE_Y, cell_k(cs[e]) / TILE_Z};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        mn[a] = min(mn[a], tc[a]);
        mx[a] = max(mx[a], tc[a]);
      }
    }
#pragma unroll
  for (int a = 0; a < 3; a++)
    for (int o = 32; o > 0; o >>= 1) {
      mn[a] = min(mn[a], __shfl_xor(mn[a], o));
      mx[a] = max(mx[a], __shfl_xor(mx[a], o));
    }
  if ((t & 63) == 0)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      red[a][t >> 6] = mn[a];
      red[3 + a][t >> 6] = mx[a];
    }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 3; a++) {
    mn[a] = red[a][0];
    mx[a] = red[3 + a][0];
    for (int q = 1; q < ST / 64; q++) {
      mn[a] = min(mn[a], red[a][q]);
      mx[a] = max(mx[a], red[3 + a][q]);
    }
    if (mx[a] < mn[a]) mn[a] = mx[a] = 0;  // no live particle
  }
  const int bx = nbits_of((uint32_t)(mx[0] - mn[0])), by = nbits_of((uint32_t)(mx[1] - mn[1]));
  // the largest relative tile any live particle can have, the dead ions'
  // key one tile past it, padding after them
  const uint32_t tmax = ((((uint32_t)(mx[2] - mn[2]) << by) | (uint32_t)(mx[1] - mn[1])) << bx) |
                        (uint32_t)(mx[0] - mn[0]);
  const uint32_t deadrel = (tmax + 1u) << 8, padrel = deadrel | 255u;
  const int nbl = nbits_of(padrel);
#pragma unroll
  for (int e = 0; e < SE; e++) {
    const uint32_t c = cs[e];
    const int i = cell_i(c), j = cell_j(c), k = cell_k(c);
    const uint32_t tr = ((((uint32_t)(k / TILE_Z - mn[2]) << by) | (uint32_t)(j / TILE_Y - mn[1]))
                         << bx) | (uint32_t)(i / TILE_X - mn[0]);
    const uint32_t local =
        (uint32_t)(i % TILE_X + TILE_X * (j % TILE_Y + TILE_Y * (k % TILE_Z)));
    key[e] = t * SE + e >= m ? padrel : c == DEAD_CELL ? deadrel : (tr << 8) | local;
  }

  // stable LSD radix sort, 4 bits per pass; thread t holds positions
  // t*SE .. t*SE+SE-1 before and after every pass
  for (int sh = 0; sh < nbl; sh += 4) {
    int d[SE], r[SE];
#pragma unroll
    for (int e = 0; e < SE; e++) d[e] = (int)((key[e] >> sh) & 15u);
#pragma unroll
    for (int b = 0; b < 16; b++) scan[pad16(b * ST + t)] = 0;
#pragma unroll
    for (int e = 0; e < SE; e++) {
      int rr = 0;
#pragma unroll
      for (int e2 = 0; e2 < e; e2++) rr += d[e2] == d[e] ? 1 : 0;
      r[e] = rr;
    }
#pragma unroll
    for (int e = 0; e < SE; e++) {
      bool last = true;
#pragma unroll
      for (int e2 = e + 1; e2 < SE; e2++) last = last && d[e2] != d[e];
      if (last) scan[pad16(d[e] * ST + t)] = (uint32_t)(r[e] + 1);
    }
    __syncthreads();
    // exclusive scan of the counters in (digit, thread) order: thread t
    // rakes entries 16t .. 16t+15
    uint32_t v[16], s = 0, tot;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      v[j] = scan[pad16(16 * t + j)];
      s += v[j];
    }
    uint32_t pre = block_excl_scan(s, wsum, &tot);
#pragma unroll
    for (int j = 0; j < 16; j++) {
      scan[pad16(16 * t + j)] = pre;
      pre += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < SE; e++) {
      int pos = (int)scan[pad16(d[e] * ST + t)] + r[e];
      skey[pos] = key[e];
      sidx[pos] = (uint16_t)idx[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < SE; e++) {
      key[e] = skey[t * SE + e];
      idx[e] = sidx[t * SE + e];
    }
  }
  if (nbl == 0) {
#pragma unroll
    for (int e = 0; e < SE; e++) skey[t * SE + e] = key[e];
  }
  __syncthreads();

  // segments: a head is a valid position whose tile differs from the
  // previous one's (the tile of a fine key k is k / TILE_CELLS)
  bool hd[SE];
  uint32_t heads = 0;
#pragma unroll
  for (int e = 0; e < SE; e++) {
    int p = t * SE + e;
    hd[e] = p < m && (p == 0 || (skey[p - 1] >> 8) != (key[e] >> 8));
    heads += hd[e] ? 1u : 0u;
  }
  uint32_t nseg;
  const uint32_t s0 = block_excl_scan(heads, wsum, &nseg);
  {
    uint32_t s = s0;
#pragma unroll
    for (int e = 0; e < SE; e++)
      if (hd[e]) scan[s++] = (uint32_t)(t * SE + e);
    if (t == 0) scan[nseg] = (uint32_t)m;
  }
  __syncthreads();
  for (uint32_t q = t; q < nseg; q += ST) {
    uint32_t st = scan[q], len = scan[q + 1] - st;
    // the segment's tile
    const uint32_t tr = skey[st] >> 8;
    uint32_t k = dead_tile;
    if (tr <= tmax) {
      const uint32_t tx = (uint32_t)mn[0] + (tr & ((1u << bx) - 1u)),
                     ty = (uint32_t)mn[1] + ((tr >> bx) & ((1u << by) - 1u)),
                     tz = (uint32_t)mn[2] + (tr >> (bx + by));
      k = tx + (uint32_t)ntx * (ty + (uint32_t)nty * tz);
    }
    segkey[c0 + q] = k;
    seginfo[c0 + q] = st | (len << 12);
    atomicAdd(cnt + k, len);
    atomicAdd(nrun + k, 1u);
  }
  int sid = (int)s0 - 1;  // a thread's leading non-heads continue the previous segment
#pragma unroll
  for (int e = 0; e < SE; e++) {
    int p = t * SE + e;
    if (hd[e]) sid++;
    if (p < m) perm[c0 + p] = idx[e] | ((uint32_t)sid << 12);
  }
  if (t == 0) nsegc[blockIdx.x] = nseg;
}

// B: each segment into its tile's list
__global__ void k_sortB(const uint32_t* nsegc, const uint32_t* segkey, const uint32_t* seginfo,
                        const uint32_t* rstart, uint32_t* rcur, uint2* runs) {
  const int64_t c0 = (int64_t)blockIdx.x * SCH;
  const uint32_t ns = nsegc[blockIdx.x];
  for (uint32_t q = threadIdx.x; q < ns; q += blockDim.x) {
    uint32_t k = segkey[c0 + q];
    uint32_t slot = rstart[k] + atomicAdd(rcur + k, 1u);
    runs[slot] = make_uint2((uint32_t)(c0 + q), seginfo[c0 + q] >> 12);
  }
}

// C: one thread per segment record: its destination is the tile's start
// plus the lengths of the tile's segments from earlier chunks
__global__ void k_sortC(const uint32_t* segkey, const uint32_t* kstart, const uint32_t* rstart,
                        int64_t nkeys, const uint2* runs, uint32_t* segbase) {
  const uint32_t nruns = rstart[nkeys];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nruns;
       q += (int64_t)gridDim.x * blockDim.x) {
    uint2 me = runs[q];
    uint32_t k = segkey[me.x];
    uint32_t base = kstart[k];
    for (uint32_t x = rstart[k]; x < rstart[k + 1]; x++) {
      uint2 o = runs[x];
      if (o.x < me.x) base += o.y;
    }
    segbase[me.x] = base;
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

// D: one workgroup per chunk walks its sorted positions; writes of a
// segment are contiguous.  Every thread resolves its DPT positions first;
// then each array of the chunk is read once, coalesced, into LDS and
// written out in sorted order from there.
constexpr int DT = 256, DPT = SCH / DT;
template <typename T>
__device__ __forceinline__ void move_array(const T* __restrict__ src, T* __restrict__ dst,
                                           T* stage, int64_t c0, int m,
                                           const uint32_t (&s)[DPT], const uint32_t (&d)[DPT],
                                           int cnt) {
  __syncthreads();  // the previous array's reads of stage are done
#pragma unroll
  for (int i = 0; i < DPT; i++) {
    int p = (int)threadIdx.x + i * DT;
    if (p < m) stage[p] = src[c0 + p];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < DPT; i++)
    if (i < cnt) dst[d[i]] = stage[s[i]];
}
template <typename R>
__global__ __launch_bounds__(DT) void k_sortD(int64_t n, const uint32_t* __restrict__ perm,
                                              const uint32_t* __restrict__ seginfo,
                                              const uint32_t* __restrict__ segbase, Move Mv) {
  __shared__ R stage[SCH];
  const int64_t c0 = (int64_t)blockIdx.x * SCH;
  const int m = (int)min((int64_t)SCH, n - c0);
  const int cnt = m > (int)threadIdx.x ? (m - (int)threadIdx.x + DT - 1) / DT : 0;
  uint32_t s[DPT], d[DPT];
#pragma unroll
  for (int i = 0; i < DPT; i++) {
    if (i < cnt) {
      int p = (int)threadIdx.x + i * DT;
      uint32_t w = perm[c0 + p];
      uint32_t sg = w >> 12;
      s[i] = w & 4095u;
      d[i] = segbase[c0 + sg] + (uint32_t)p - (seginfo[c0 + sg] & 4095u);
    }
  }
  move_array(Mv.cell, Mv.cell2, (uint32_t*)stage, c0, m, s, d, cnt);
#pragma unroll
  for (int a = 0; a < 9; a++)
    move_array((const R*)Mv.p[a], (R*)Mv.p2[a], stage, c0, m, s, d, cnt);
  if (Mv.bp) move_array(Mv.bp, Mv.bp2, (float*)stage, c0, m, s, d, cnt);
  if (Mv.mr) move_array(Mv.mr, Mv.mr2, (float*)stage, c0, m, s, d, cnt);
  if (Mv.id) move_array(Mv.id, Mv.id2, (uint32_t*)stage, c0, m, s, d, cnt);
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
// box of an item is its tile's extent
__global__ void k_items_count(int64_t ntiles, const uint32_t* start, uint32_t* nit) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t a = start[t], b = start[t + 1];
  nit[t] = (b - a + ITEM_MAX - 1) / ITEM_MAX;
}
__global__ void k_items_fill(int64_t ntiles, const uint32_t* start, const uint32_t* nit,
                             const uint32_t* ioff, int ntx, int nty, int ncx, int ncy, int ncz,
                             Item* items, int* boxes) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t a = start[t], b = start[t + 1];
  int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), tz = (int)(t / ((int64_t)ntx * nty));
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
    bx[2] = tz * TILE_Z;
    bx[3] = min((tx + 1) * TILE_X, ncx) - 1;
    bx[4] = min((ty + 1) * TILE_Y, ncy) - 1;
    bx[5] = min((tz + 1) * TILE_Z, ncz) - 1;
  }
}

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

// exclusive scan of n counts into out (n+1 entries: out[n] = total); bsum
// holds n / 4096 + 1 block sums (the context's sblk: up to nkeys + 1 counts)
int scan(swap_ctx* c, int64_t n, const uint32_t* in, uint32_t* out, uint32_t* bsum = nullptr) {
  if (!bsum) bsum = c->sblk;
  int64_t nb = (n + SCAN_T * SCAN_E - 1) / (SCAN_T * SCAN_E);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SCAN_T), 0, c->stream, n, in, out,
                     bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, c->stream, nb, bsum, out + n);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, out, bsum);
  return post(c);
}

}  // namespace

int so_scan_u32(swap_ctx* c, int64_t n, const uint32_t* in, uint32_t* out, uint32_t* bsum) {
  return scan(c, n, in, out, bsum);
}

int so_sort(swap_ctx* c) {
  int64_t n = c->n;
  const int64_t nk = c->nkeys, nt = c->ntiles;
  // an open x run has one more tile, nt, for the particles that left; they
  // end up behind the live ones, which are the first sstart[nt]
  const int64_t nkd = nt + (c->g.ox ? 1 : 0);
  const int ntx = c->ntile[0], nty = c->ntile[1];
  if (n > 0) {
    const int64_t nch = (n + SCH - 1) / SCH;
    HIPCHK(c, hipMemsetAsync(c->scnt, 0, sizeof(uint32_t) * nkd, c->stream));
    HIPCHK(c, hipMemsetAsync(c->snrun, 0, sizeof(uint32_t) * nkd, c->stream));
    hipLaunchKernelGGL(k_sortA, dim3((unsigned)nch), dim3(ST), 0, c->stream, n, c->cell, ntx,
                       nty, (uint32_t)nt, c->scnt, c->snrun, c->segkey, c->seginfo, c->nsegc,
                       c->srbase);
    if (int e = post(c)) return e;
    if (int e = scan(c, nkd, c->scnt, c->sstart)) return e;
    if (int e = scan(c, nkd, c->snrun, c->srstart)) return e;
    HIPCHK(c, hipMemsetAsync(c->scur, 0, sizeof(uint32_t) * nkd, c->stream));
    hipLaunchKernelGGL(k_sortB, dim3((unsigned)nch), dim3(256), 0, c->stream, c->nsegc,
                       c->segkey, c->seginfo, c->srstart, c->scur, c->sruns);
    hipLaunchKernelGGL(k_sortC, dim3(8192), dim3(256), 0, c->stream, c->segkey, c->sstart,
                       c->srstart, nkd, c->sruns, c->segbase);
    Move Mv;
    Mv.cell = c->cell; Mv.cell2 = c->cell2;
    for (int a = 0; a < 9; a++) { Mv.p[a] = c->p[a]; Mv.p2[a] = c->p2[a]; }
    Mv.bp = c->bp; Mv.bp2 = c->bp2; Mv.mr = c->mr; Mv.mr2 = c->mr2;
    Mv.id = c->id; Mv.id2 = c->id2;
    if (c->prec == 64)
      hipLaunchKernelGGL(k_sortD<double>, dim3((unsigned)nch), dim3(256), 0, c->stream, n,
                         c->srbase, c->seginfo, c->segbase, Mv);
    else
      hipLaunchKernelGGL(k_sortD<float>, dim3((unsigned)nch), dim3(256), 0, c->stream, n,
                         c->srbase, c->seginfo, c->segbase, Mv);
    if (int e = post(c)) return e;
    std::swap(c->cell, c->cell2);
    for (int a = 0; a < 9; a++) std::swap(c->p[a], c->p2[a]);
    std::swap(c->bp, c->bp2);
    std::swap(c->mr, c->mr2);
    std::swap(c->id, c->id2);
    if (c->g.ox) {
      uint32_t live = 0;
      HIPCHK(c, hipMemcpyAsync(&live, c->sstart + nt, 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      c->escaped += n - (int64_t)live;
      c->dead_pending = false;
      // the ions that left (now behind the live ones) leave the input account
      if (int e = sr_input_add(c, live, n - (int64_t)live, -1.0)) return e;
      c->n = n = live;
    }
  } else {
    HIPCHK(c, hipMemsetAsync(c->sstart, 0, sizeof(uint32_t) * (nt + 1), c->stream));
  }
  // work items and their cell boxes; the item count stays on the device
  uint32_t* nit = c->scur;  // reuse: ntiles + 1 entries each
  uint32_t* ioff = c->scur + c->ntiles + 1;
  hipLaunchKernelGGL(k_items_count, dim3(nblk(c->ntiles, 256)), dim3(256), 0, c->stream,
                     c->ntiles, c->sstart, nit);
  if (int e = scan(c, c->ntiles, nit, ioff)) return e;
  hipLaunchKernelGGL(k_items_fill, dim3(nblk(c->ntiles, 256)), dim3(256), 0, c->stream, c->ntiles,
                     c->sstart, nit, ioff, ntx, nty, c->cfg.ncx, c->cfg.ncy, c->cfg.ncz,
                     c->items, c->boxes);
  HIPCHK(c, hipMemcpyAsync(c->d_nitems, ioff + c->ntiles, sizeof(int), hipMemcpyDeviceToDevice,
                           c->stream));
  if (int e = post(c)) return e;
  c->n_sorted = n;
  c->steps_since_sort = 0;
  c->sorts++;
  return 0;
}
