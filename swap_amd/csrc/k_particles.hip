// k_particles.hip -- routine-level particle kernels, one per reference routine
// of gutsp.f.  One thread per particle; they materialise what the reference
// materialises (ijkp, wght, Ep, vplus, vminus) and deposit with compensated
// global atomics.  The fused passes of swap_step are in k_pass.hip.
#include "particle_math.h"

using namespace pm;

namespace {

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
// exact (fixed-point) global atomics with the ghost fold applied to the node
// index (gutsp.f:1480-1498, 1704-1805)
// stride > 1 deposits only the particles with id % stride == 0, each with
// stride times its weight (a low-ppc subsample of the same particles);
// stride < -1 deposits the complement, id % |stride| != 0, each with
// |stride| / (|stride| - 1) times its weight (a sample disjoint from it).
// back != 0 deposits at x - (dt/2) v, the position before the last half move
// (where swap_step's final deposit took place).
template <typename R>
__global__ void k_deposit_simple(Geo g, int64_t n, PDev<R> P, int mode,
                                 typename AccOf<R>::T* acc, const uint32_t* id, int stride,
                                 PCoef<R> K, int back) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (stride > 1 && id[p] % (uint32_t)stride != 0u) return;
  if (stride < -1 && id[p] % (uint32_t)(-stride) == 0u) return;
  uint32_t c = P.cell[p];
  int ci = cell_i(c), cj = cell_j(c), ck = cell_k(c);
  R f0 = P.fx[p], f1 = P.fy[p], f2 = P.fz[p];
  if (back) {
    ci = wrapc(ci + half_move(f0, -P.vx[p], K.s[0]), K.nc[0]);
    cj = wrapc(cj + half_move(f1, -P.vy[p], K.s[1]), K.nc[1]);
    ck = wrapc(ck + half_move(f2, -P.vz[p], K.s[2]), K.nc[2]);
  }
  R w[8];
  weights8(f0, f1, f2, w);
  R r = P.bp ? (R)1 / (R)P.bp[p] : (R)1;
  if (stride > 1) r *= (R)stride;
  if (stride < -1) r *= (R)(-stride) / (R)(-stride - 1);
  R q[3] = {0, 0, 0};
  if (mode == 1) { q[0] = P.vx[p]; q[1] = P.vy[p]; q[2] = P.vz[p]; }
  else if (mode == 2) {
    q[0] = (R)1.5 * P.vx[p] - (R)0.5 * P.v1x[p];
    q[1] = (R)1.5 * P.vy[p] - (R)0.5 * P.v1y[p];
    q[2] = (R)1.5 * P.vz[p] - (R)0.5 * P.v1z[p];
  } else if (mode == 3) { q[0] = P.v1x[p]; q[1] = P.v1y[p]; q[2] = P.v1z[p]; }
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    int u = ci + (cc & 1), v = cj + ((cc >> 2) & 1), t = ck + ((cc >> 1) & 1);
    int64_t node = gidx(g, node_of_x(g, u), node_of(v, g.ncy), node_of(t, g.ncz));
    R wr = w[cc] * r;
    if (mode == 0) acc_add(acc, node, wr, fix_bits(0));
    else
      for (int m = 0; m < 3; m++)
        acc_add(acc, node + (int64_t)(m + 1) * g.nn, wr * q[m], fix_bits(m + 1));
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

// improve_up's vp1 = (v+ + v-)/2 (gutsp.f:420-423)
template <typename R>
__global__ void k_improve(int64_t n, int64_t cap, PDev<R> P, const R* vpl, const R* vmi) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  P.v1x[p] = (R)0.5 * (vpl[p] + vmi[p]);
  P.v1y[p] = (R)0.5 * (vpl[p + cap] + vmi[p + cap]);
  P.v1z[p] = (R)0.5 * (vpl[p + 2 * cap] + vmi[p + 2 * cap]);
}

// get_vp_final (gutsp.f:443-446)
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

// move_ion_half (gutsp.f:454-510) on split positions, periodic in every axis
template <typename R>
__global__ void k_move(int64_t n, PDev<R> P, PCoef<R> K) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  R f0 = P.fx[p], f1 = P.fy[p], f2 = P.fz[p];
  int i = wrapc(cell_i(c) + half_move(f0, P.vx[p], K.s[0]), K.nc[0]);
  int j = wrapc(cell_j(c) + half_move(f1, P.vy[p], K.s[1]), K.nc[1]);
  int k = wrapc(cell_k(c) + half_move(f2, P.vz[p], K.s[2]), K.nc[2]);
  P.fx[p] = f0; P.fy[p] = f1; P.fz[p] = f2;
  P.cell[p] = pack_cell(i, j, k);
}

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

int ensure_routine_scratch(swap_ctx* c) {
  if (c->Ep) return 0;
  size_t b3 = c->rsz * 3 * (size_t)c->cap;
  HIPCHK(c, hipMalloc(&c->Ep, b3 ? b3 : 16));
  HIPCHK(c, hipMalloc(&c->vpl, b3 ? b3 : 16));
  HIPCHK(c, hipMalloc(&c->vmi, b3 ? b3 : 16));
  HIPCHK(c, hipMalloc(&c->wght, c->rsz * 8 * (size_t)c->cap + 16));
  HIPCHK(c, hipMalloc(&c->ijkp, sizeof(int32_t) * 3 * (size_t)c->cap + 16));
  return 0;
}

}  // namespace

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

#define PL(kern, n, ...) \
  hipLaunchKernelGGL(kern, dim3(nblk((n), 256)), dim3(256), 0, c->stream, __VA_ARGS__)

int pl_ensure_scratch(swap_ctx* c) { return ensure_routine_scratch(c); }

int pl_weights(swap_ctx* c) {
  if (int e = ensure_routine_scratch(c)) return e;
  if (c->n == 0) return 0;
  PDISPATCH(c, PL((k_weights<R>), c->n, c->n, c->cap, pdev<R>(c), c->ijkp, (R*)c->wght));
  return post(c);
}
int pl_deposit_simple(swap_ctx* c, int mode, int stride, bool back) {
  if (c->n == 0) return 0;
  if ((stride > 1 || stride < -1) && !c->id)
    return set_err(c, SWAP_ESTATE, "subsampled deposit needs ids");
  PDISPATCH(c, PL((k_deposit_simple<R>), c->n, c->g, c->n, pdev<R>(c), mode,
                  (typename AccOf<R>::T*)c->macc, c->id, stride, pcoef<R>(c), back ? 1 : 0));
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
