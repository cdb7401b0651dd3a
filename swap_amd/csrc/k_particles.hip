// k_particles.hip -- particle kernels of the hybrid step.
//
// Particle state is structure-of-arrays with split-precision positions: a
// packed 0-based cell index (i | j<<11 | k<<21) plus an in-cell fraction
// f in (0,1] per axis.  Cells are right-closed like the reference's search
// (q(i) < x <= q(i+1), gutsp.f:1391-1412), so x = q(i) + f dx.  The eight
// trilinear weights of get_interp_weights (gutsp.f:1415-1429) are products
// of f and 1-f.
//
// Two families of kernels:
//  * routine-level kernels, one thread per particle, one per reference
//    routine; they materialise what the reference materialises (ijkp, wght,
//    Ep, vplus, vminus) and deposit with global compensated atomics;
//  * the fused passes of swap_step: P1 moments at x^n, P2 the u_p predictor,
//    P3 the final push, both half moves and the x^{n+1/2} moments.  A pass
//    runs one workgroup per work item (a run of one tile's particles after
//    the counting sort); each lane walks PASS_KP consecutive particles and
//    folds same-cell runs in registers, runs are flushed into a compensated
//    FP32 LDS window covering the item's deposit footprint, and the window
//    is flushed to HBM once per item.
#include "swap_internal.h"

namespace {

template <typename R>
__device__ __forceinline__ void weights8(R fx, R fy, R fz, R w[8]) {
  R gx = (R)1 - fx, gy = (R)1 - fy, gz = (R)1 - fz;
  R a00 = gx * gy, a10 = fx * gy, a01 = gx * fy, a11 = fx * fy;
  w[0] = a00 * gz; w[1] = a10 * gz; w[2] = a00 * fz; w[3] = a10 * fz;
  w[4] = a01 * gz; w[5] = a11 * gz; w[6] = a01 * fz; w[7] = a11 * fz;
}

// storage offsets of the 8 corners in the reference's weight order
__device__ __forceinline__ void corner_off(const Geo& g, int64_t o[8]) {
  int64_t sx = 1, sy = g.nx, sz = (int64_t)g.nx * g.ny;
  o[0] = 0; o[1] = sx; o[2] = sz; o[3] = sx + sz;
  o[4] = sy; o[5] = sx + sy; o[6] = sy + sz; o[7] = sx + sy + sz;
}

// trilinear gather of the 6-component source S = (ajc-upc, btc)
// (gutsp.f:217-251); returns A = gathered (aj-up), B = gathered btc
template <typename R>
__device__ __forceinline__ void gather6(const Geo& g, const R* __restrict__ S, int64_t base,
                                        const R w[8], R A[3], R B[3]) {
  int64_t off[8];
  corner_off(g, off);
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const R* s = S + (int64_t)m * g.nn + base;
    R v = s[off[0]] * w[0] + s[off[1]] * w[1] + s[off[2]] * w[2] + s[off[3]] * w[3] +
          s[off[4]] * w[4] + s[off[5]] * w[5] + s[off[6]] * w[6] + s[off[7]] * w[7];
    if (m < 3) A[m] = v; else B[m - 3] = v;
  }
}

// Ep = ((aj - up) x B) mrat, gutsp.f:253-271
template <typename R>
__device__ __forceinline__ void ep_of(const R A[3], const R B[3], R mr, R E[3]) {
  E[0] = (A[1] * B[2] - A[2] * B[1]) * mr;
  E[1] = (A[2] * B[0] - A[0] * B[2]) * mr;
  E[2] = (A[0] * B[1] - A[1] * B[0]) * mr;
}

// v- = v + dt/2 Ep and the Crank-Nicolson rotation, gutsp.f:302-397
template <typename R>
__device__ __forceinline__ void rotate(R dt, const R v[3], const R E[3], const R B[3], R mr,
                                       R vm[3], R vp[3]) {
  R hdt = (R)0.5 * dt;
#pragma unroll
  for (int m = 0; m < 3; m++) vm[m] = v[m] + hdt * E[m];
  R vxb0 = vm[1] * B[2] * mr - vm[2] * B[1] * mr;
  R vxb1 = vm[2] * B[0] * mr - vm[0] * B[2] * mr;
  R vxb2 = vm[0] * B[1] * mr - vm[1] * B[0] * mr;
  R vdb = vm[0] * B[0] * mr + vm[1] * B[1] * mr + vm[2] * B[2] * mr;
  R Bx = B[0] * mr, By = B[1] * mr, Bz = B[2] * mr;
  R B2 = Bx * Bx + By * By + Bz * Bz;
  R dt2 = dt * dt;
  R ad = (R)1.0 / ((R)1.0 + ((R)0.25 * B2 * dt2));
  R a1 = ((R)1 - ((R)0.25 * B2 * dt2)) * ad;
  R a2 = dt * ad;
  R a3 = (R)0.5 * dt2 * ad;
  vp[0] = a1 * vm[0] + a2 * vxb0 + a3 * vdb * B[0] * mr;
  vp[1] = a1 * vm[1] + a2 * vxb1 + a3 * vdb * B[1] * mr;
  vp[2] = a1 * vm[2] + a2 * vxb2 + a3 * vdb * B[2] * mr;
}

// split-precision half move along one axis (move_ion_half, gutsp.f:471-510):
// t = f + v dt/(2 d) in cell units, integer part carried into the cell,
// fraction kept in (0,1].  ceil() gives the right-closed convention exactly.
__device__ __forceinline__ int half_move(float& f, float v, float s) {
  float t = fmaf(v, s, f);
  float sh = ceilf(t) - 1.0f;
  f = t - sh;
  return (int)sh;
}
__device__ __forceinline__ int half_move(double& f, double v, double s) {
  double t = fma(v, s, f);
  double sh = ceil(t) - 1.0;
  f = t - sh;
  return (int)sh;
}
__device__ __forceinline__ int wrapc(int c, int nc) {
  c %= nc;
  return c < 0 ? c + nc : c;
}

// ---------------- routine-level kernels ----------------

template <typename R>
__global__ void k_weights(int64_t n, int64_t cap, PDev<R> P, int32_t* ijkp, R* wght) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  ijkp[p] = cell_i(c) + 1;
  ijkp[p + cap] = cell_j(c) + 1;
  ijkp[p + 2 * cap] = cell_k(c) + 1;
  R w[8];
  weights8(P.fx[p], P.fy[p], P.fz[p], w);
#pragma unroll
  for (int q = 0; q < 8; q++) wght[p + q * cap] = w[q];
}

// update_np (mode 0) / update_up (mode 1: vp, 2: 1.5 vp - 0.5 vp1, 3: vp1),
// one thread per particle, compensated global atomics with the ghost fold
// applied to the node index (gutsp.f:1480-1498, 1704-1805)
template <typename R>
__global__ void k_deposit_simple(Geo g, int64_t n, PDev<R> P, int mode, R* hi, R* lo) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  int ci = cell_i(c), cj = cell_j(c), ck = cell_k(c);
  R w[8];
  weights8(P.fx[p], P.fy[p], P.fz[p], w);
  R r = P.bp ? (R)1 / (R)P.bp[p] : (R)1;
  R q[3];
  if (mode == 1) { q[0] = P.vx[p]; q[1] = P.vy[p]; q[2] = P.vz[p]; }
  else if (mode == 2) {
    q[0] = (R)1.5 * P.vx[p] - (R)0.5 * P.v1x[p];
    q[1] = (R)1.5 * P.vy[p] - (R)0.5 * P.v1y[p];
    q[2] = (R)1.5 * P.vz[p] - (R)0.5 * P.v1z[p];
  } else if (mode == 3) { q[0] = P.v1x[p]; q[1] = P.v1y[p]; q[2] = P.v1z[p]; }
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    int u = ci + (cc & 1), v = cj + ((cc >> 2) & 1), t = ck + ((cc >> 1) & 1);
    int64_t node = gidx(g, node_of(u, g.ncx), node_of(v, g.ncy), node_of(t, g.ncz));
    R wr = w[cc] * r;
    if (mode == 0) acc_add(hi, lo, node, wr);
    else
      for (int m = 0; m < 3; m++) acc_add(hi, lo, node + (int64_t)(m + 1) * g.nn, wr * q[m]);
  }
}

template <typename R>
__global__ void k_get_Ep(Geo g, int64_t n, int64_t cap, PDev<R> P, const R* S, R* Ep) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  R w[8], A[3], B[3], E[3];
  weights8(P.fx[p], P.fy[p], P.fz[p], w);
  gather6(g, S, gidx(g, cell_i(c), cell_j(c), cell_k(c)), w, A, B);
  ep_of(A, B, P.mr ? (R)P.mr[p] : (R)1, E);
  Ep[p] = E[0]; Ep[p + cap] = E[1]; Ep[p + 2 * cap] = E[2];
}

// get_vplus_vminus gathers btc itself (gutsp.f:359-369): B from the field
template <typename R>
__global__ void k_vpm(Geo g, int64_t n, int64_t cap, PDev<R> P, R dt, const R* btc, const R* Ep,
                      R* vpl, R* vmi) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  R w[8];
  weights8(P.fx[p], P.fy[p], P.fz[p], w);
  int64_t base = gidx(g, cell_i(c), cell_j(c), cell_k(c)), off[8];
  corner_off(g, off);
  R B[3];
  for (int m = 0; m < 3; m++) {
    const R* s = btc + (int64_t)m * g.nn + base;
    B[m] = s[off[0]] * w[0] + s[off[1]] * w[1] + s[off[2]] * w[2] + s[off[3]] * w[3] +
           s[off[4]] * w[4] + s[off[5]] * w[5] + s[off[6]] * w[6] + s[off[7]] * w[7];
  }
  R v[3] = {P.vx[p], P.vy[p], P.vz[p]};
  R E[3] = {Ep[p], Ep[p + cap], Ep[p + 2 * cap]};
  R vm[3], vp[3];
  rotate(dt, v, E, B, P.mr ? (R)P.mr[p] : (R)1, vm, vp);
  for (int m = 0; m < 3; m++) {
    vmi[p + m * cap] = vm[m];
    vpl[p + m * cap] = vp[m];
  }
}

template <typename R>
__global__ void k_improve(int64_t n, int64_t cap, PDev<R> P, const R* vpl, const R* vmi) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  P.v1x[p] = (R)0.5 * (vpl[p] + vmi[p]);
  P.v1y[p] = (R)0.5 * (vpl[p + cap] + vmi[p + cap]);
  P.v1z[p] = (R)0.5 * (vpl[p + 2 * cap] + vmi[p + 2 * cap]);
}

template <typename R>
__global__ void k_final(int64_t n, int64_t cap, PDev<R> P, R dt, const R* Ep, const R* vpl) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  R hdt = (R)0.5 * dt;
  P.v1x[p] = P.vx[p]; P.v1y[p] = P.vy[p]; P.v1z[p] = P.vz[p];
  P.vx[p] = vpl[p] + hdt * Ep[p];
  P.vy[p] = vpl[p + cap] + hdt * Ep[p + cap];
  P.vz[p] = vpl[p + 2 * cap] + hdt * Ep[p + 2 * cap];
}

template <typename R>
__global__ void k_move(int64_t n, PDev<R> P, PCoef<R> K) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  if (c == 0xFFFFFFFFu) return;
  R f0 = P.fx[p], f1 = P.fy[p], f2 = P.fz[p];
  int i = wrapc(cell_i(c) + half_move(f0, P.vx[p], K.s[0]), K.nc[0]);
  int j = wrapc(cell_j(c) + half_move(f1, P.vy[p], K.s[1]), K.nc[1]);
  int k = wrapc(cell_k(c) + half_move(f2, P.vz[p], K.s[2]), K.nc[2]);
  P.fx[p] = f0; P.fy[p] = f1; P.fz[p] = f2;
  P.cell[p] = pack_cell(i, j, k);
}

// ---------------- fused passes ----------------

__device__ __forceinline__ void lds_add(float* hi, float* lo, int i, float v) {
  if (v == 0.f) return;
  float old = atomicAdd(hi + i, v);
  float e = two_sum_err(old, v, old + v);
  if (e != 0.f) atomicAdd(lo + i, e);
}
__device__ __forceinline__ void lds_add(double* hi, double* lo, int i, double v) {
  (void)lo;
  if (v == 0.0) return;
  atomicAdd(hi + i, v);
}
__device__ __forceinline__ void glob_add2(float* hi, float* lo, int64_t i, float h, float l) {
  if (h != 0.f) {
    float old = atomicAdd(hi + i, h);
    l += two_sum_err(old, h, old + h);
  }
  if (l != 0.f) atomicAdd(lo + i, l);
}
__device__ __forceinline__ void glob_add2(double* hi, double* lo, int64_t i, double h, double l) {
  (void)lo;
  (void)l;
  if (h != 0.0) atomicAdd(hi + i, h);
}

// nearest periodic image of cell c to a box centre
__device__ __forceinline__ int unwrap(int c, int centre, int nc) {
  int d = c - centre;
  if (2 * d > nc) c -= nc;
  else if (2 * d < -nc) c += nc;
  return c;
}

template <int MODE>
struct NMom {
  static constexpr int value = MODE == 2 ? 3 : 4;
};

// MODE 1: moments at x^n of r and r(1.5 v - 0.5 v1)          (maind.f:385-399)
// MODE 2: Ep, v+-, deposit r (v+ + v-)/2 at x^n               (maind.f:400-402)
// MODE 3: Ep, v+-, v1 <- v, v <- v+ + dt/2 Ep, move to x^{n+1/2}, deposit r
//         and r v there, move to x^{n+1}                     (maind.f:404-413, 479)
template <typename R, int MODE>
__global__ __launch_bounds__(PASS_WG) void k_pass(Geo g, PDev<R> P, PCoef<R> K,
                                                  const R* __restrict__ S, R* __restrict__ ghi,
                                                  R* __restrict__ glo, const Item* __restrict__ items,
                                                  const int* __restrict__ d_nitems,
                                                  int64_t n_sorted, int64_t n_total,
                                                  const int* __restrict__ boxes, int* boxes_next,
                                                  int mx, int my, int mz,
                                                  unsigned long long* dcnt) {
  constexpr int M = NMom<MODE>::value;
  constexpr bool F32 = sizeof(R) == 4;
  __shared__ R lhi[WIN_VALUES];
  __shared__ R llo[F32 ? WIN_VALUES : 1];
  __shared__ int sbox[8];

  const int tid = threadIdx.x;
  const int nitems = *d_nitems;
  int64_t i0, i1;
  int bl[3] = {0, 0, 0}, bh[3] = {-1, -1, -1};
  bool windowed = false;
  int w = blockIdx.x;
  if (w < nitems) {
    Item it = items[w];
    i0 = it.start;
    i1 = it.end;
    const int* b = boxes + 6 * w;
    int nc3[3] = {g.ncx, g.ncy, g.ncz};
    for (int a = 0; a < 3; a++) {
      bl[a] = b[a];
      bh[a] = b[3 + a];
      // keep the box centre in [0, nc)
      int cen = (bl[a] + bh[a]) >> 1;
      int sh = (cen >= 0 ? cen / nc3[a] : -((-cen + nc3[a] - 1) / nc3[a])) * nc3[a];
      bl[a] -= sh;
      bh[a] -= sh;
    }
    windowed = bh[0] >= bl[0];
  } else {
    int64_t t = w - nitems;
    i0 = n_sorted + t * ITEM_MAX;
    if (i0 >= n_total) return;
    i1 = i0 + ITEM_MAX < n_total ? i0 + ITEM_MAX : n_total;
  }
  const int marg[3] = {MODE == 3 ? mx : 0, MODE == 3 ? my : 0, MODE == 3 ? mz : 0};
  int u0[3], W[3];
  int cen[3];
  const int nc[3] = {g.ncx, g.ncy, g.ncz};
  for (int a = 0; a < 3; a++) {
    u0[a] = bl[a] - marg[a];
    W[a] = bh[a] - bl[a] + 2 + 2 * marg[a];
    cen[a] = (bl[a] + bh[a]) >> 1;
  }
  const int Wn = W[0] * W[1] * W[2];
  if (windowed && (int64_t)Wn * M > WIN_VALUES) windowed = false;
  if (windowed) {
    for (int q = tid; q < Wn * M; q += PASS_WG) {
      lhi[q] = (R)0;
      if (F32) llo[q] = (R)0;
    }
  }
  if (MODE == 3 && tid < 6) sbox[tid] = tid < 3 ? 0x7fffffff : -0x7fffffff;
  __syncthreads();

  const int coff[8] = {0, 1, W[0] * W[1], 1 + W[0] * W[1], W[0], 1 + W[0],
                       W[0] + W[0] * W[1], 1 + W[0] + W[0] * W[1]};
  R acc[8][M];
  int rkey = -1;
  int bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int bmax[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
  unsigned long long misses = 0;

  auto flush = [&]() {
    if (rkey >= 0) {
#pragma unroll
      for (int cc = 0; cc < 8; cc++)
#pragma unroll
        for (int m = 0; m < M; m++) lds_add(lhi, llo, rkey + coff[cc] + m * Wn, acc[cc][m]);
    }
    rkey = -1;
  };
  // deposit one particle at unwrapped cell (ux,uy,uz), weights w, moments q
  auto deposit = [&](int ux, int uy, int uz, const R wt[8], const R q[M]) {
    int lx = ux - u0[0], ly = uy - u0[1], lz = uz - u0[2];
    bool in = windowed && lx >= 0 && ly >= 0 && lz >= 0 && lx < W[0] - 1 && ly < W[1] - 1 &&
              lz < W[2] - 1;
    if (in) {
      int key = lx + W[0] * (ly + W[1] * lz);
      if (key != rkey) {
        flush();
        rkey = key;
#pragma unroll
        for (int cc = 0; cc < 8; cc++)
#pragma unroll
          for (int m = 0; m < M; m++) acc[cc][m] = wt[cc] * q[m];
      } else {
#pragma unroll
        for (int cc = 0; cc < 8; cc++)
#pragma unroll
          for (int m = 0; m < M; m++) acc[cc][m] += wt[cc] * q[m];
      }
    } else {
      misses++;
#pragma unroll
      for (int cc = 0; cc < 8; cc++) {
        int64_t node = gidx(g, node_of(ux + (cc & 1), g.ncx), node_of(uy + ((cc >> 2) & 1), g.ncy),
                            node_of(uz + ((cc >> 1) & 1), g.ncz));
#pragma unroll
        for (int m = 0; m < M; m++) {
          R v = wt[cc] * q[m];
          if (v != (R)0) glob_add2(ghi, glo, node + (int64_t)m * g.nn, v, (R)0);
        }
      }
    }
  };

  for (int64_t base = i0 + (int64_t)tid * PASS_KP; base < i1; base += (int64_t)PASS_WG * PASS_KP) {
    uint32_t cl[PASS_KP];
    R fx[PASS_KP], fy[PASS_KP], fz[PASS_KP], vx[PASS_KP], vy[PASS_KP], vz[PASS_KP];
    R ex[MODE == 1 ? PASS_KP : 1], ey[MODE == 1 ? PASS_KP : 1], ez[MODE == 1 ? PASS_KP : 1];
#pragma unroll
    for (int t = 0; t < PASS_KP; t++) {
      int64_t p = base + t;
      bool ok = p < i1;
      cl[t] = ok ? P.cell[p] : 0xFFFFFFFFu;
      fx[t] = ok ? P.fx[p] : (R)0;
      fy[t] = ok ? P.fy[p] : (R)0;
      fz[t] = ok ? P.fz[p] : (R)0;
      vx[t] = ok ? P.vx[p] : (R)0;
      vy[t] = ok ? P.vy[p] : (R)0;
      vz[t] = ok ? P.vz[p] : (R)0;
      if constexpr (MODE == 1) {
        ex[t] = ok ? P.v1x[p] : (R)0;
        ey[t] = ok ? P.v1y[p] : (R)0;
        ez[t] = ok ? P.v1z[p] : (R)0;
      }
    }
#pragma unroll
    for (int t = 0; t < PASS_KP; t++) {
      uint32_t c = cl[t];
      if (c == 0xFFFFFFFFu) continue;
      int64_t p = base + t;
      int ci = cell_i(c), cj = cell_j(c), ck = cell_k(c);
      R r = P.bp ? (R)1 / (R)P.bp[p] : (R)1;
      R w8[8];
      weights8(fx[t], fy[t], fz[t], w8);
      int ux = unwrap(ci, cen[0], g.ncx), uy = unwrap(cj, cen[1], g.ncy),
          uz = unwrap(ck, cen[2], g.ncz);
      if constexpr (MODE == 1) {
        R q[M];
        q[0] = r;
        q[1] = r * ((R)1.5 * vx[t] - (R)0.5 * ex[t]);
        q[2] = r * ((R)1.5 * vy[t] - (R)0.5 * ey[t]);
        q[3] = r * ((R)1.5 * vz[t] - (R)0.5 * ez[t]);
        deposit(ux, uy, uz, w8, q);
      } else {
        R A[3], B[3], E[3], vm[3], vpl[3];
        gather6(g, S, gidx(g, ci, cj, ck), w8, A, B);
        R mr = P.mr ? (R)P.mr[p] : (R)1;
        ep_of(A, B, mr, E);
        R v[3] = {vx[t], vy[t], vz[t]};
        rotate(K.dt, v, E, B, mr, vm, vpl);
        if constexpr (MODE == 2) {
          R q[M];
#pragma unroll
          for (int m = 0; m < 3; m++) q[m] = r * ((R)0.5 * (vpl[m] + vm[m]));
          deposit(ux, uy, uz, w8, q);
        } else {
          R hdt = (R)0.5 * K.dt;
          R vn[3];
#pragma unroll
          for (int m = 0; m < 3; m++) vn[m] = vpl[m] + hdt * E[m];
          P.v1x[p] = vx[t]; P.v1y[p] = vy[t]; P.v1z[p] = vz[t];
          P.vx[p] = vn[0]; P.vy[p] = vn[1]; P.vz[p] = vn[2];
          R f0 = fx[t], f1 = fy[t], f2 = fz[t];
          ux += half_move(f0, vn[0], K.s[0]);
          uy += half_move(f1, vn[1], K.s[1]);
          uz += half_move(f2, vn[2], K.s[2]);
          {
            R wh[8];
            weights8(f0, f1, f2, wh);
            R q[M] = {r, r * vn[0], r * vn[1], r * vn[2]};
            deposit(ux, uy, uz, wh, q);
          }
          ux += half_move(f0, vn[0], K.s[0]);
          uy += half_move(f1, vn[1], K.s[1]);
          uz += half_move(f2, vn[2], K.s[2]);
          P.fx[p] = f0; P.fy[p] = f1; P.fz[p] = f2;
          P.cell[p] = pack_cell(wrapc(ux, g.ncx), wrapc(uy, g.ncy), wrapc(uz, g.ncz));
          bmin[0] = min(bmin[0], ux); bmax[0] = max(bmax[0], ux);
          bmin[1] = min(bmin[1], uy); bmax[1] = max(bmax[1], uy);
          bmin[2] = min(bmin[2], uz); bmax[2] = max(bmax[2], uz);
        }
      }
    }
  }
  flush();
  if (misses) atomicAdd(dcnt, misses);
  if (MODE == 3) {
    for (int a = 0; a < 3; a++) {
      atomicMin(&sbox[a], bmin[a]);
      atomicMax(&sbox[3 + a], bmax[a]);
    }
  }
  __syncthreads();
  if (MODE == 3 && w < nitems && tid < 6) {
    int* bn = boxes_next + 6 * w;
    if (tid < 3) atomicMin(bn + tid, sbox[tid]);
    else atomicMax(bn + tid, sbox[tid]);
  }
  if (windowed) {
    for (int q = tid; q < Wn; q += PASS_WG) {
      int lx = q % W[0], ly = (q / W[0]) % W[1], lz = q / (W[0] * W[1]);
      int64_t node = gidx(g, node_of(u0[0] + lx, g.ncx), node_of(u0[1] + ly, g.ncy),
                          node_of(u0[2] + lz, g.ncz));
#pragma unroll
      for (int m = 0; m < M; m++) {
        R h = lhi[q + m * Wn];
        R l = F32 ? llo[q + m * Wn] : (R)0;
        if (h != (R)0 || l != (R)0) glob_add2(ghi, glo, node + (int64_t)m * g.nn, h, l);
      }
    }
  }
}

__global__ void k_box_reset(int64_t n, int* b) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n * 6) return;
  b[q] = (q % 6) < 3 ? 0x7fffffff : -0x7fffffff;
}

#define PDISPATCH(c, ...)   \
  do {                      \
    if ((c)->prec == 64) {  \
      using R = double;     \
      __VA_ARGS__;          \
    } else {                \
      using R = float;      \
      __VA_ARGS__;          \
    }                       \
  } while (0)

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

int ensure_routine_scratch(swap_ctx* c) {
  if (c->Ep) return 0;
  size_t b3 = c->rsz * 3 * (size_t)c->cap;
  HIPCHK(c, hipMalloc(&c->Ep, b3));
  HIPCHK(c, hipMalloc(&c->vpl, b3));
  HIPCHK(c, hipMalloc(&c->vmi, b3));
  HIPCHK(c, hipMalloc(&c->wght, c->rsz * 8 * (size_t)c->cap));
  HIPCHK(c, hipMalloc(&c->ijkp, sizeof(int32_t) * 3 * (size_t)c->cap));
  return 0;
}

}  // namespace

#define PL(kern, n, ...) \
  hipLaunchKernelGGL(kern, dim3(nblk((n), 256)), dim3(256), 0, c->stream, __VA_ARGS__)

int pl_weights(swap_ctx* c) {
  if (int e = ensure_routine_scratch(c)) return e;
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_weights<R>), c->n, c->n, c->cap, pdev<R>(c), c->ijkp, (R*)c->wght));
  return post(c);
}
int pl_deposit_simple(swap_ctx* c, int mode) {
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_deposit_simple<R>), c->n, c->g, c->n, pdev<R>(c), mode, (R*)c->mhi,
                  (R*)c->mlo));
  return post(c);
}
int pl_get_Ep(swap_ctx* c) {
  if (int e = ensure_routine_scratch(c)) return e;
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_get_Ep<R>), c->n, c->g, c->n, c->cap, pdev<R>(c), (const R*)c->S,
                  (R*)c->Ep));
  return post(c);
}
int pl_vplus_vminus(swap_ctx* c) {
  if (int e = ensure_routine_scratch(c)) return e;
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_vpm<R>), c->n, c->g, c->n, c->cap, pdev<R>(c), (R)c->cfg.dt,
                  (const R*)c->fld[SWAP_F_BTC], (const R*)c->Ep, (R*)c->vpl, (R*)c->vmi));
  return post(c);
}
int pl_improve_vp1(swap_ctx* c) {
  if (int e = ensure_routine_scratch(c)) return e;
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_improve<R>), c->n, c->n, c->cap, pdev<R>(c), (const R*)c->vpl,
                  (const R*)c->vmi));
  return post(c);
}
int pl_vp_final(swap_ctx* c) {
  if (int e = ensure_routine_scratch(c)) return e;
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_final<R>), c->n, c->n, c->cap, pdev<R>(c), (R)c->cfg.dt, (const R*)c->Ep,
                  (const R*)c->vpl));
  return post(c);
}
int pl_move_half(swap_ctx* c) {
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_move<R>), c->n, c->n, pdev<R>(c), pcoef<R>(c)));
  return post(c);
}

int pl_box_reset(swap_ctx* c) {
  hipLaunchKernelGGL(k_box_reset, dim3(nblk(c->max_items * 6, 256)), dim3(256), 0, c->stream,
                     c->max_items, c->boxes_next);
  return post(c);
}

int pl_pass(swap_ctx* c, int mode) {
  if (c->n == 0) return 0;
  int64_t ns = c->n_sorted < 0 ? 0 : c->n_sorted;
  int64_t tail = c->n - ns;
  int64_t nb = (c->n_sorted < 0 ? 0 : c->max_items) + (tail + ITEM_MAX - 1) / ITEM_MAX;
  if (nb == 0) return 0;
  dim3 grid((unsigned)nb), blk(PASS_WG);
  int mx = c->margin[0], my = c->margin[1], mz = c->margin[2];
  int zero = 0;
  const int* dn = c->n_sorted < 0 ? (const int*)c->dred : c->d_nitems;
  if (c->n_sorted < 0) HIPCHK(c, hipMemcpyAsync(c->dred, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
  PDISPATCH(c, {
    PDev<R> P = pdev<R>(c);
    PCoef<R> K = pcoef<R>(c);
    if (mode == 1)
      hipLaunchKernelGGL((k_pass<R, 1>), grid, blk, 0, c->stream, c->g, P, K, (const R*)c->S,
                         (R*)c->mhi, (R*)c->mlo, c->items, dn, ns, c->n, c->boxes,
                         c->boxes_next, mx, my, mz, c->dcnt);
    else if (mode == 2)
      hipLaunchKernelGGL((k_pass<R, 2>), grid, blk, 0, c->stream, c->g, P, K, (const R*)c->S,
                         (R*)c->mhi + c->g.nn, (R*)c->mlo + c->g.nn, c->items, dn, ns, c->n,
                         c->boxes, c->boxes_next, mx, my, mz, c->dcnt);
    else
      hipLaunchKernelGGL((k_pass<R, 3>), grid, blk, 0, c->stream, c->g, P, K, (const R*)c->S,
                         (R*)c->mhi, (R*)c->mlo, c->items, dn, ns, c->n, c->boxes,
                         c->boxes_next, mx, my, mz, c->dcnt);
  });
  return post(c);
}
