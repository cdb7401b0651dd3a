// particle_math.h -- per-particle arithmetic shared by the routine-level and
// the fused particle kernels.
//
// Particle state is structure-of-arrays with split-precision positions: a
// packed 0-based cell index (i | j<<11 | k<<21) plus an in-cell fraction
// f in (0,1] per axis.  Cells are right-closed like the reference's search
// (q(i) < x <= q(i+1), gutsp.f:1391-1412), so x = q(i) + f dx and the eight
// trilinear weights of get_interp_weights (gutsp.f:1415-1429) are products
// of f and 1-f.
#pragma once
#include "swap_internal.h"

namespace pm {

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

// The gather source S = (ajc-upc, btc) is node-major: SREC values per node
// (6 used, 2 zero, swap_internal.h), so a corner is two 16-byte loads in FP32.

// the 6 source values of node q (node index, not scaled by SREC)
template <typename R>
__device__ __forceinline__ void src_node(const R* __restrict__ S, int64_t q, R v[6]) {
  const R* p = S + q * SREC;
#pragma unroll
  for (int m = 0; m < 6; m++) v[m] = p[m];
}
template <>
__device__ __forceinline__ void src_node<float>(const float* __restrict__ S, int64_t q,
                                                float v[6]) {
  const float4 a = *(const float4*)(S + q * SREC);
  const float2 b = *(const float2*)(S + q * SREC + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
}

// trilinear gather of S (gutsp.f:217-251); A = gathered (aj - up), B =
// gathered btc; the corners are summed in the reference's corner order.
// Addresses: one record pointer per y-z corner pair (the x neighbour is the
// next record, an immediate offset), so the 8 corners take 4 pointers.
template <typename R>
__device__ __forceinline__ void gather6(const Geo& g, const R* __restrict__ S, int64_t base,
                                        const R w[8], R A[3], R B[3]) {
  const int64_t sy = (int64_t)g.nx * SREC, sz = (int64_t)g.nx * g.ny * SREC;
  const R* p0 = S + base * SREC;
  const R* pc[4] = {p0, p0 + sz, p0 + sy, p0 + sy + sz};  // corners 0-1, 2-3, 4-5, 6-7
  R v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 8; k++) {
    R c[6];
    src_node(pc[k >> 1], (int64_t)(k & 1), c);
#pragma unroll
    for (int m = 0; m < 6; m++) v[m] += c[m] * w[k];
  }
#pragma unroll
  for (int m = 0; m < 3; m++) {
    A[m] = v[m];
    B[m] = v[3 + m];
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
// fraction kept in (0,1]; ceil() gives the right-closed convention exactly
// and t - (ceil(t) - 1) is exact, so the fraction never leaves (0,1].
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
__device__ __forceinline__ int wrapc(int c, int nc) { return wrap_range(c, nc); }

// order-preserving int image of a float (for integer atomic min/max)
__device__ __forceinline__ int fkey(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float fval(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

}  // namespace pm
