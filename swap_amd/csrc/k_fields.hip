// k_fields.hip -- staggered-grid field kernels of the hybrid step.
//
// Placement (the reference's Yee layout): n, ct and node averages at nodes;
// aj, up, E component m on the m-directed edge; B component m on the
// m-normal face.  Every kernel evaluates its stencil at the canonical node of
// each storage node, so ghost layers come out equal to what boundary.f's
// periodic() would copy (boundary.f:9-52) and no separate ghost pass runs.
// B is held and advanced in FP64; everything else in the context precision R.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "swap_internal.h"

namespace {

struct FC {  // uniform-grid constants
  double dx, dy, dz;
  double rx, ry, rz;  // xrat, yrat, zrat: 0.5 on a uniform grid (initial.f:276)
  double alpha;
};

FC fc_of(swap_ctx* c) {
  FC f;
  f.dx = c->cfg.dx; f.dy = c->cfg.dy; f.dz = c->cfg.dz;
  f.rx = f.ry = f.rz = 0.5;
  f.alpha = c->cfg.alpha;
  return f;
}

#define GRID_LOOP_HEAD                                                     \
  int i = blockIdx.x * blockDim.x + threadIdx.x;                           \
  int j = blockIdx.y * blockDim.y + threadIdx.y;                           \
  int k = blockIdx.z;                                                      \
  if (i >= g.nx || j >= g.ny) return;                                      \
  int ci = canon_x(g, i), cj = canon(j, g.ny), ck = canon(k, g.nz);        \
  const int64_t o = gidx(g, i, j, k);                                      \
  (void)o;                                                                 \
  (void)ci; (void)cj; (void)ck;

#define AT(arr, m, ii, jj, kk) (arr)[gidx(g, ii, jj, kk) + (int64_t)(m)*g.nn]

dim3 gblock() { return dim3(64, 4, 1); }
dim3 ggrid(const Geo& g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }

// B component m at face (i,j,k); with b2 given, the half-level average
// btp1 = (b1p2 + b1)/2 of get_Ep1 (gutsf.f:1210-1219)
__device__ __forceinline__ double bval(const Geo& g, const double* b, const double* b2, int m,
                                       int i, int j, int k) {
  int64_t q = gidx(g, i, j, k) + (int64_t)m * g.nn;
  return b2 ? 0.5 * (b[q] + b2[q]) : b[q];
}

// edge_to_center stencil (grid_interp.f:22-55) on a face-centred field
template <typename T>
__device__ __forceinline__ void e2c3(const Geo& g, const FC& f, const T* b, const T* b2, int i,
                                     int j, int k, double o3[3]) {
  auto B = [&](int m, int ii, int jj, int kk) -> double {
    int64_t q = gidx(g, ii, jj, kk) + (int64_t)m * g.nn;
    return b2 ? 0.5 * ((double)b[q] + (double)b2[q]) : (double)b[q];
  };
  double hi, lo;
  hi = B(0, i, j - 1, k) + f.ry * (B(0, i, j, k) - B(0, i, j - 1, k));
  lo = B(0, i, j - 1, k - 1) + f.ry * (B(0, i, j, k - 1) - B(0, i, j - 1, k - 1));
  o3[0] = lo + f.rz * (hi - lo);
  hi = B(1, i - 1, j, k) + f.rx * (B(1, i, j, k) - B(1, i - 1, j, k));
  lo = B(1, i - 1, j, k - 1) + f.rx * (B(1, i, j, k - 1) - B(1, i - 1, j, k - 1));
  o3[1] = lo + f.rz * (hi - lo);
  hi = B(2, i, j - 1, k) + f.ry * (B(2, i, j, k) - B(2, i, j - 1, k));
  lo = B(2, i - 1, j - 1, k) + f.ry * (B(2, i - 1, j, k) - B(2, i - 1, j - 1, k));
  o3[2] = lo + f.rx * (hi - lo);
}

// face_to_center stencil (grid_interp.f:150-170) on an edge-centred field
template <typename R>
__device__ __forceinline__ void f2c3(const Geo& g, const FC& f, const R* v, int i, int j, int k,
                                     R o3[3]) {
  o3[0] = (R)f.rx * (AT(v, 0, i, j, k) - AT(v, 0, i - 1, j, k)) + AT(v, 0, i - 1, j, k);
  o3[1] = (R)f.ry * (AT(v, 1, i, j, k) - AT(v, 1, i, j - 1, k)) + AT(v, 1, i, j - 1, k);
  o3[2] = (R)f.rz * (AT(v, 2, i, j, k) - AT(v, 2, i, j, k - 1)) + AT(v, 2, i, j, k - 1);
}

// curlB, gutsf.f:178-244: aj = curl B / (alpha n_edge); optional a = aj - up
template <typename R>
__global__ void k_curlB(Geo g, FC f, const double* b, const double* b2, const R* np, const R* up,
                        R* aj, R* a) {
  GRID_LOOP_HEAD
  R n0 = np[gidx(g, ci, cj, ck)];
  R nt1 = (R)0.5 * (n0 + np[gidx(g, ci + 1, cj, ck)]);
  R nt2 = (R)0.5 * (n0 + np[gidx(g, ci, cj + 1, ck)]);
  R nt3 = (R)0.5 * (n0 + np[gidx(g, ci, cj, ck + 1)]);
  double c1 = (bval(g, b, b2, 2, ci, cj, ck) - bval(g, b, b2, 2, ci, cj - 1, ck)) / f.dy +
              (bval(g, b, b2, 1, ci, cj, ck - 1) - bval(g, b, b2, 1, ci, cj, ck)) / f.dz;
  double c2 = (bval(g, b, b2, 0, ci, cj, ck) - bval(g, b, b2, 0, ci, cj, ck - 1)) / f.dz +
              (bval(g, b, b2, 2, ci - 1, cj, ck) - bval(g, b, b2, 2, ci, cj, ck)) / f.dx;
  double c3 = (bval(g, b, b2, 1, ci, cj, ck) - bval(g, b, b2, 1, ci - 1, cj, ck)) / f.dx +
              (bval(g, b, b2, 0, ci, cj - 1, ck) - bval(g, b, b2, 0, ci, cj, ck)) / f.dy;
  R al = (R)f.alpha;
  R j1 = (R)c1 / (nt1 * al), j2 = (R)c2 / (nt2 * al), j3 = (R)c3 / (nt3 * al);
  aj[o] = j1; aj[o + g.nn] = j2; aj[o + 2 * g.nn] = j3;
  if (a) {
    int64_t oc = gidx(g, ci, cj, ck);
    a[o] = j1 - up[oc];
    a[o + g.nn] = j2 - up[oc + g.nn];
    a[o + 2 * g.nn] = j3 - up[oc + 2 * g.nn];
  }
}

// get_E's node part (gutsf.f:1067-1069): ct = face_to_center(a) x edge_to_center(B)
template <typename R>
__global__ void k_cross_nodes(Geo g, FC f, const R* a, const double* b, R* ct) {
  GRID_LOOP_HEAD
  R aa[3];
  f2c3(g, f, a, ci, cj, ck, aa);
  double bd[3];
  e2c3<double>(g, f, b, nullptr, ci, cj, ck, bd);
  R bx = (R)bd[0], by = (R)bd[1], bz = (R)bd[2];
  ct[o] = aa[1] * bz - aa[2] * by;
  ct[o + g.nn] = aa[2] * bx - aa[0] * bz;
  ct[o + 2 * g.nn] = aa[0] * by - aa[1] * bx;
}

// edge_to_center of B (optionally of (b+b2)/2), into a real node field
template <typename R>
__global__ void k_e2c_B(Geo g, FC f, const double* b, const double* b2, R* out) {
  GRID_LOOP_HEAD
  double bd[3];
  e2c3<double>(g, f, b, b2, ci, cj, ck, bd);
  out[o] = (R)bd[0]; out[o + g.nn] = (R)bd[1]; out[o + 2 * g.nn] = (R)bd[2];
}

// edge_to_center of a real field (routine-level API)
template <typename R>
__global__ void k_e2c(Geo g, FC f, const R* in, R* out) {
  GRID_LOOP_HEAD
  double bd[3];
  e2c3<R>(g, f, in, nullptr, ci, cj, ck, bd);
  out[o] = (R)bd[0]; out[o + g.nn] = (R)bd[1]; out[o + 2 * g.nn] = (R)bd[2];
}

// get_Ep1's node part with the smoothed B (gutsf.f:1266-1273):
// btmf = edge_to_face average of btc0 (grid_interp.f:119-127), then
// face_to_center(btmf), crossed with face_to_center(a)
template <typename R>
__global__ void k_cross_smooth(Geo g, FC f, const R* a, const R* btc0, R* ct) {
  GRID_LOOP_HEAD
  R aa[3];
  f2c3(g, f, a, ci, cj, ck, aa);
  // the face average below node 1 is the average's own ghost: periodic, or
  // on an open x a zero-gradient copy of the average at node 1
  int im = canon_x(g, ci - 1), jm = canon(cj - 1, g.ny), km = canon(ck - 1, g.nz);
  R m1c = (R)0.5 * (AT(btc0, 0, ci, cj, ck) + AT(btc0, 0, ci + 1, cj, ck));
  R m1m = (R)0.5 * (AT(btc0, 0, im, cj, ck) + AT(btc0, 0, im + 1, cj, ck));
  R m2c = (R)0.5 * (AT(btc0, 1, ci, cj, ck) + AT(btc0, 1, ci, cj + 1, ck));
  R m2m = (R)0.5 * (AT(btc0, 1, ci, jm, ck) + AT(btc0, 1, ci, jm + 1, ck));
  R m3c = (R)0.5 * (AT(btc0, 2, ci, cj, ck) + AT(btc0, 2, ci, cj, ck + 1));
  R m3m = (R)0.5 * (AT(btc0, 2, ci, cj, km) + AT(btc0, 2, ci, cj, km + 1));
  R bx = (R)f.rx * (m1c - m1m) + m1m;
  R by = (R)f.ry * (m2c - m2m) + m2m;
  R bz = (R)f.rz * (m3c - m3m) + m3m;
  ct[o] = aa[1] * bz - aa[2] * by;
  ct[o + g.nn] = aa[2] * bx - aa[0] * bz;
  ct[o + 2 * g.nn] = aa[0] * by - aa[1] * bx;
}

// crossf2's edge average plus E = c + nu aj (gutsf.f:84-98, 1078-1085)
template <typename R>
__global__ void k_edgeavg_E(Geo g, const R* ct, const R* aj, const R* nu, R* E) {
  GRID_LOOP_HEAD
  int64_t oc = gidx(g, ci, cj, ck);
  R nuv = nu[oc];
  E[o] = (R)0.5 * (AT(ct, 0, ci, cj, ck) + AT(ct, 0, ci + 1, cj, ck)) + nuv * aj[oc];
  E[o + g.nn] = (R)0.5 * (AT(ct, 1, ci, cj, ck) + AT(ct, 1, ci, cj + 1, ck)) + nuv * aj[oc + g.nn];
  E[o + 2 * g.nn] =
      (R)0.5 * (AT(ct, 2, ci, cj, ck) + AT(ct, 2, ci, cj, ck + 1)) + nuv * aj[oc + 2 * g.nn];
}

// curlE (gutsf.f:262-303) in FP64
template <typename R>
__device__ __forceinline__ void curlE3(const Geo& g, const FC& f, const R* E, int i, int j, int k,
                                       double c[3]) {
  c[0] = ((double)AT(E, 2, i, j + 1, k) - (double)AT(E, 2, i, j, k)) / f.dy +
         ((double)AT(E, 1, i, j, k) - (double)AT(E, 1, i, j, k + 1)) / f.dz;
  c[1] = ((double)AT(E, 2, i, j, k) - (double)AT(E, 2, i + 1, j, k)) / f.dx +
         ((double)AT(E, 0, i, j, k + 1) - (double)AT(E, 0, i, j, k)) / f.dz;
  c[2] = ((double)AT(E, 0, i, j, k) - (double)AT(E, 0, i, j + 1, k)) / f.dy +
         ((double)AT(E, 1, i + 1, j, k) - (double)AT(E, 1, i, j, k)) / f.dx;
}

// predict_B / correct_B update (gutsf.f:1158-1159, 1349-1350):
// out = base - coef * curlE(E), coef = 2 dtsub or dtsub
template <typename R>
__global__ void k_curlE_update(Geo g, FC f, const R* E, const double* base, double coef,
                               double* out) {
  GRID_LOOP_HEAD
  double cE[3];
  curlE3(g, f, E, ci, cj, ck, cE);
  int64_t oc = gidx(g, ci, cj, ck);
#pragma unroll
  for (int m = 0; m < 3; m++) out[o + m * g.nn] = base[oc + m * g.nn] - coef * cE[m];
}

template <typename R>
__global__ void k_curlE(Geo g, FC f, const R* E, R* out) {
  GRID_LOOP_HEAD
  double cE[3];
  curlE3(g, f, E, ci, cj, ck, cE);
#pragma unroll
  for (int m = 0; m < 3; m++) out[o + m * g.nn] = (R)cE[m];
}

template <typename R>
__global__ void k_f2c(Geo g, FC f, const R* in, R* out) {
  GRID_LOOP_HEAD
  R v[3];
  f2c3(g, f, in, ci, cj, ck, v);
  out[o] = v[0]; out[o + g.nn] = v[1]; out[o + 2 * g.nn] = v[2];
}

// second half of edge_to_face (grid_interp.f:119-127) from node values btc0
template <typename R>
__global__ void k_face_avg(Geo g, const R* btc0, R* out) {
  GRID_LOOP_HEAD
  out[o] = (R)0.5 * (AT(btc0, 0, ci, cj, ck) + AT(btc0, 0, ci + 1, cj, ck));
  out[o + g.nn] = (R)0.5 * (AT(btc0, 1, ci, cj, ck) + AT(btc0, 1, ci, cj + 1, ck));
  out[o + 2 * g.nn] = (R)0.5 * (AT(btc0, 2, ci, cj, ck) + AT(btc0, 2, ci, cj, ck + 1));
}

// crossf2 (gutsf.f:40-104) for the routine-level API
template <typename R>
__global__ void k_crossf2(Geo g, const R* aa, const R* bb, R* out) {
  GRID_LOOP_HEAD
  auto cr = [&](int m, int ii, int jj, int kk) -> R {
    R ax = AT(aa, 0, ii, jj, kk), ay = AT(aa, 1, ii, jj, kk), az = AT(aa, 2, ii, jj, kk);
    R bx = AT(bb, 0, ii, jj, kk), by = AT(bb, 1, ii, jj, kk), bz = AT(bb, 2, ii, jj, kk);
    return m == 0 ? ay * bz - az * by : (m == 1 ? az * bx - ax * bz : ax * by - ay * bx);
  };
  out[o] = (R)0.5 * (cr(0, ci, cj, ck) + cr(0, ci + 1, cj, ck));
  out[o + g.nn] = (R)0.5 * (cr(1, ci, cj, ck) + cr(1, ci, cj + 1, ck));
  out[o + 2 * g.nn] = (R)0.5 * (cr(2, ci, cj, ck) + cr(2, ci, cj, ck + 1));
}

template <typename T>
__global__ void k_periodic(Geo g, T* f, int ncomp) {
  GRID_LOOP_HEAD
  if (ci == i && cj == j && ck == k) return;
  int64_t oc = gidx(g, ci, cj, ck);
  for (int m = 0; m < ncomp; m++) f[o + m * g.nn] = f[oc + m * g.nn];
}

template <typename R>
__global__ void k_a_from(Geo g, const R* aj, const R* up, R* a) {
  GRID_LOOP_HEAD
  int64_t oc = gidx(g, ci, cj, ck);
#pragma unroll
  for (int m = 0; m < 3; m++) a[o + m * g.nn] = aj[oc + m * g.nn] - up[oc + m * g.nn];
}

// one node record of the gather source (6 values; the 2 pad values stay 0)
template <typename R>
__device__ __forceinline__ void put_src(R* p, const R v[6]) {
#pragma unroll
  for (int m = 0; m < 6; m++) p[m] = v[m];
}
template <>
__device__ __forceinline__ void put_src<float>(float* p, const float v[6]) {
  *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
  *(float2*)(p + 4) = make_float2(v[4], v[5]);
}

// get_Ep's node fields (gutsp.f:183-184): ajc = face_to_center(aj),
// upc = face_to_center(up); the particle gather source is (ajc-upc, btc)
template <typename R>
__global__ void k_build_src(Geo g, FC f, const R* aj, const R* up, const R* btc, R* ajc, R* upc,
                            R* S) {
  GRID_LOOP_HEAD
  R jc[3], uc[3];
  f2c3(g, f, aj, ci, cj, ck, jc);
  f2c3(g, f, up, ci, cj, ck, uc);
  int64_t oc = gidx(g, ci, cj, ck);
  R rec[6];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    ajc[o + m * g.nn] = jc[m];
    if (upc) upc[o + m * g.nn] = uc[m];
    rec[m] = jc[m] - uc[m];
    rec[3 + m] = btc[oc + m * g.nn];
  }
  put_src(S + o * SREC, rec);
}

template <typename R>
__global__ void k_update_src(Geo g, FC f, const R* up, const R* ajc, R* upc, R* S) {
  GRID_LOOP_HEAD
  R uc[3];
  f2c3(g, f, up, ci, cj, ck, uc);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    if (upc) upc[o + m * g.nn] = uc[m];
    S[o * SREC + m] = ajc[o + m * g.nn] - uc[m];
  }
}

// np = deposit sum / (vol beta) at every node (update_np's scaling,
// gutsp.f:1480-1491; the fold of :1496-1498 happened at deposit time)
template <typename R>
__global__ void k_finalize_np(Geo g, const typename AccOf<R>::T* a, R qn, R* np) {
  GRID_LOOP_HEAD
  double v = acc_read(a, gidx(g, ci, cj, ck), fix_bits(0)) * (double)qn;
  if (g.ox) v *= face_scale(g, ci);
  np[o] = (R)v;
}

// update_up's division by np and node->edge average (gutsp.f:1704-1756,
// 1825-1836); a points at moment 1 (x velocity), moments are nn apart
template <typename R>
__global__ void k_finalize_up(Geo g, const typename AccOf<R>::T* a, const R* np, R qn, R* up) {
  GRID_LOOP_HEAD
  int64_t q0 = gidx(g, ci, cj, ck);
  const int ix = canon_x(g, ci + 1);
  int64_t qn1[3] = {gidx(g, ix, cj, ck), gidx(g, ci, canon(cj + 1, g.ny), ck),
                    gidx(g, ci, cj, canon(ck + 1, g.nz))};
  R n0 = np[q0];
  // np of an open x face node carries its share factor; the velocity sums
  // take the same factor so that up is their plain ratio
  const R s0 = (R)face_scale(g, ci), s1x = (R)face_scale(g, ix);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    R n1 = np[qn1[m]];
    R a0 = (R)acc_read(a, q0 + m * g.nn, fix_bits(1)) * qn;
    R a1 = (R)acc_read(a, qn1[m] + m * g.nn, fix_bits(1)) * qn;
    if (g.ox) {
      a0 *= s0;
      a1 *= m == 0 ? s1x : s0;
    }
    R c0 = n0 != (R)0 ? a0 / n0 : (R)0;
    R c1 = n1 != (R)0 ? a1 / n1 : (R)0;
    up[o + m * g.nn] = (R)0.5 * (c0 + c1);
  }
}

// check_time_step's per-node whistler bound (gutsf.f:1782-1792)
// A grid-stride loop over all storage nodes with a bounded number of
// workgroups; the grid minimum takes one atomic per workgroup (atomics on
// one address serialise).  Non-negative floats order as their bit patterns.
constexpr int DT_BLOCKS = 1024;
template <typename R>
__global__ __launch_bounds__(256) void k_deltat(Geo g, FC f, const double* b, const R* np,
                                                R* dtn, float* dmin) {
  __shared__ unsigned int wmin[4];
  unsigned int u = 0xFFFFFFFFu;
  for (int64_t o = blockIdx.x * (int64_t)256 + threadIdx.x; o < g.nn;
       o += (int64_t)gridDim.x * 256) {
    R b1 = (R)b[o], b2 = (R)b[o + g.nn], b3 = (R)b[o + 2 * g.nn];
    R ak = (R)2. / (R)f.dx;
    R btot = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
    R al = (R)f.alpha;
    R a1 = ak * ak * btot / (al * np[o]);
    R a2 = (ak * btot) * (ak * btot) / (al * np[o]);
    R womega = (R)0.5 * (a1 + sqrt(a1 * a1 + (R)4 * a2));
    R phi = womega / ak;
    R dt = (R)f.dx / phi;
    dtn[o] = dt;
    float d = (float)dt;
    if (!(d >= 0.f)) d = 0.f;  // NaN or negative: treat as failing
    u = min(u, __float_as_uint(d));
  }
  for (int off = 32; off > 0; off >>= 1) u = min(u, (unsigned int)__shfl_xor((int)u, off));
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = u;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMin((unsigned int*)dmin, min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3])));
}

// check_time_step's halving (gutsf.f:1793-1796) on the device: if the grid
// minimum of deltat fails the first test, the reference halves dtsub once per
// failing node, in its loop order (i outer, j, k inner) and against the dtsub
// current at that node; one workgroup walks the nodes in that order (a rare
// path).  Writes dtsub (in the context precision, as the host's restatement
// rounds it) and ntf = ntsub 2^h to ts[0], ts[1], where the subcycle kernels
// and the host read them.
template <typename R>
__global__ void k_tstep(Geo g, const R* dtn, const float* dmin, double dtsub0, double thr64,
                        float thr32, int ntsub, double* ts) {
  __shared__ int flag[1024];
  const float mn = *dmin;
  const bool maybe = sizeof(R) == 8 ? ((double)mn <= thr64) : (mn <= thr32);
  R dtsub = (R)dtsub0;
  int h = 0;
  if (maybe) {
    const int64_t total = g.nn;
    for (int64_t base = 0; base < total; base += 1024) {
      int64_t r = base + threadIdx.x;
      int cand = 0;
      if (r < total) {
        int k = (int)(r % g.nz);
        int j = (int)((r / g.nz) % g.ny);
        int i = (int)(r / ((int64_t)g.nz * g.ny));
        cand = dtn[gidx(g, i, j, k)] <= (R)2.0 * (R)dtsub0;
      }
      flag[threadIdx.x] = cand;
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int t = 0; t < 1024 && base + t < total; t++) {
          if (!flag[t]) continue;
          int64_t rr = base + t;
          int k = (int)(rr % g.nz);
          int j = (int)((rr / g.nz) % g.ny);
          int i = (int)(rr / ((int64_t)g.nz * g.ny));
          if (dtn[gidx(g, i, j, k)] <= (R)2.0 * dtsub) {
            dtsub = dtsub / (R)2.0;
            h++;
          }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    ts[0] = (double)dtsub;
    ts[1] = (double)ntsub * (double)(1ll << (h < 62 ? h : 62));
  }
}
__global__ void k_fill_f32(float* p, float v) { *p = v; }
// dtsub given by the caller (swap_field_subcycles)
__global__ void k_set_dtsub(double* ts, double dtsub, double ntf) {
  ts[0] = dtsub;
  ts[1] = ntf;
}

// ---------------- fused, LDS-tiled field subcycle ----------------
//
// One subcycle (maind.f:420-476) in two launches instead of nine.
//   predictor: aj = curlB(b1)/(alpha n), a = aj - up; ct = face_to_center(a)
//     x edge_to_center(b1); E = edge average of ct + nu aj; b1p2 = b12 -
//     2 dtsub curlE(E)                                   (gutsf.f:1115-1169);
//     it stores bh = (b1p2 + b1)/2, all the corrector reads of b1p2 (b1 of
//     its output node comes from its own LDS ring), so that the corrector
//     stages one FP64 level instead of two
//   corrector: the same chain on bh with get_Ep1's smoothed B
//     (edge_to_center, then node->face->centre averages); b = b1 - dtsub
//     curlE(E)                                 (gutsf.f:1174-1304, 1309-1359)
// A workgroup owns TX x TY columns of canonical nodes and a run of z planes.
// It marches up z keeping a ring of planes of every intermediate in LDS, each
// with the x/y halo its consumer's stencil needs, so the inputs are read once
// per tile (halo re-reads are L2 hits) and no intermediate goes to HBM.  The
// arithmetic is that of the per-routine kernels above, term by term, in the
// same precisions.  Results are written to every storage image of a
// canonical node (ghosts included), so the fields stay periodic.
//
// LDS plane layout: tile position (x, y) relative to the origin (X0, Y0),
// x in [-2, TX+2], y in [-2, TY+2], at index (x+2) + PW (y+2); each position
// holds the value of the canonical node it wraps to.

// canonical index [1, n-2] of any (unwrapped) storage coordinate s
__device__ __forceinline__ int wrapc(int s, int n) {
  const int m = n - 2;
  int r = (s - 1) % m;
  if (r < 0) r += m;
  return r + 1;
}

// storage indices that hold canonical index c along an axis of n nodes:
// periodic, or (open, an inflow/outflow x) the zero-gradient ghost beside
// each face node
__device__ __forceinline__ int images(int c, int n, int s[2], bool open = false) {
  s[0] = c;
  if (c == 1) { s[1] = open ? 0 : n - 1; return 2; }
  if (c == n - 2) { s[1] = open ? n - 1 : 0; return 2; }
  return 1;
}

// LDS tile x position -> canonical x node: periodic, or on an open x the
// clamp to [1, ncx] (the node whose copy a zero-gradient ghost holds)
__device__ __forceinline__ int tile_cx(const Geo& g, int s) {
  return g.ox ? (s < 1 ? 1 : (s > g.ncx ? g.ncx : s)) : wrapc(s, g.nx);
}

// An open x: after a phase has written an intermediate over the tile, the
// positions of x ghosts (storage x < 1 or > ncx) take the value of the node
// they clamp to -- what the per-routine kernels leave in a ghost layer, so
// that the next phase's stencils read the same numbers.  a: component 0 of
// ncomp planes PA apart; positions x0 - 2 + (t % PW) along x.
template <int PW, int PA, int NT, typename T>
__device__ __forceinline__ void ghost_x(T* a, int ncomp, int tid, int x0, int ncx) {
  for (int t = tid; t < PA; t += NT) {
    const int xp = t % PW, sx = x0 - 2 + xp;
    if (sx >= 1 && sx <= ncx) continue;
    const int src = t - xp + (sx < 1 ? 1 : ncx) - x0 + 2;
    for (int m = 0; m < ncomp; m++) a[t + m * PA] = a[src + m * PA];
  }
}

template <typename T>
__device__ __forceinline__ void put3(const Geo& g, T* dst, int x, int y, int z, T v0, T v1,
                                     T v2) {
  int xs[2], ys[2], zs[2];
  const int ax = images(x, g.nx, xs, g.ox), ay = images(y, g.ny, ys), az = images(z, g.nz, zs);
  for (int c = 0; c < az; c++)
    for (int b = 0; b < ay; b++)
      for (int a = 0; a < ax; a++) {
        const int64_t o = gidx(g, xs[a], ys[b], zs[c]);
        dst[o] = v0;
        dst[o + g.nn] = v1;
        dst[o + 2 * g.nn] = v2;
      }
}

template <int TX, int TY>
struct FT {
  static constexpr int PW = TX + 5, PH = TY + 5, PA = PW * PH, NT = TX * TY;
};

// 3-slot ring index of plane z (z may be negative)
__device__ __forceinline__ int r3(int z) { return ((z % 3) + 3) % 3; }

// the positions a thread owns in one phase: t = tid + k NT over an N-position
// region W wide whose first position sits at LDS (ox, oy).  p[k] = LDS index
// (-1: none), q[k] = offset of the canonical node within its z plane.
template <int K, int NT, int PW>
__device__ __forceinline__ void own(int tid, int W, int N, int ox, int oy, const int* tcx,
                                    const int* tcy, int nx, int p[K], int q[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int t = tid + k * NT;
    p[k] = t < N ? (t % W + ox) + PW * (t / W + oy) : -1;
    q[k] = t < N ? tcx[p[k] % PW] + nx * tcy[p[k] / PW] : 0;
  }
}

// the predictor's LDS rings
template <typename R, int TX, int TY>
struct PredLds {
  static constexpr int PW = FT<TX, TY>::PW, PH = FT<TX, TY>::PH, PA = FT<TX, TY>::PA;
  double sb[2][3][PA];  // b1, planes s+1, s+2
  R sn[2][PA];          // np, planes s+2, s+3
  R sa[2][3][PA];       // a = aj - up
  R sj[2][3][PA];       // aj
  R sc[2][3][PA];       // ct
  R se[2][3][PA];       // E
  int tcx[PW], tcy[PH];
};

// the predictor over tile (bx, by, bz): TX x TY columns, z planes
// [1 + bz tz, 1 + (bz + 1) tz); coef = 2 dtsub
template <typename R, int TX, int TY, int NT>
__device__ __forceinline__ void sub_predict_tile(const Geo& g, const FC& f,
                                                 const double* __restrict__ b1,
                                                 const double* __restrict__ b12,
                                                 const R* __restrict__ np,
                                                 const R* __restrict__ up,
                                                 const R* __restrict__ nu, double coef, int bx,
                                                 int by, int bz, int tz,
                                                 double* __restrict__ out,
                                                 PredLds<R, TX, TY>& L) {
  constexpr int PW = FT<TX, TY>::PW, PH = FT<TX, TY>::PH, PA = FT<TX, TY>::PA;
  static_assert(NT >= TX * TY, "one thread per output column at least");
  auto& sb = L.sb;
  auto& sn = L.sn;
  auto& sa = L.sa;
  auto& sj = L.sj;
  auto& sc = L.sc;
  auto& se = L.se;
  auto& tcx = L.tcx;
  auto& tcy = L.tcy;
  const int tid = threadIdx.x;
  const int X0 = 1 + bx * TX, Y0 = 1 + by * TY, Z0 = 1 + bz * tz;
  const int Z1 = min(Z0 + tz, g.ncz + 1);
  if (Z0 > g.ncz) return;  // empty z chunk (whole workgroup)
  for (int t = tid; t < PW; t += NT) tcx[t] = tile_cx(g, X0 - 2 + t);
  for (int t = tid; t < PH; t += NT) tcy[t] = wrapc(Y0 - 2 + t, g.ny);
  __syncthreads();
  const int64_t nxy = (int64_t)g.nx * g.ny, nn = g.nn;
  const R al = (R)f.alpha;
  // an open x and a tile whose halo reaches an x ghost (uniform per workgroup)
  const bool xface = g.ox && (X0 - 2 < 1 || X0 + TX + 2 > g.ncx);
  // positions per phase: the whole plane (staging), curl [-1, T+1], E [0, T]
  constexpr int AW = TX + 3, AN = AW * (TY + 3), EW = TX + 1, EN = EW * (TY + 1);
  constexpr int KL = (PA + NT - 1) / NT, KA = (AN + NT - 1) / NT, KE = (EN + NT - 1) / NT;
  int lp[KL], lq[KL], ap[KA], aq[KA], ep[KE], eq[KE];
  own<KL, NT, PW>(tid, PW, PA, 0, 0, tcx, tcy, g.nx, lp, lq);
  own<KA, NT, PW>(tid, AW, AN, 1, 1, tcx, tcy, g.nx, ap, aq);
  own<KE, NT, PW>(tid, EW, EN, 2, 2, tcx, tcy, g.nx, ep, eq);
  const int olx = tid % TX, oly = tid / TX, ox = X0 + olx, oy = Y0 + oly;
  const bool oown = tid < TX * TY && ox <= g.ncx && oy <= g.ncy;
  // the global data of step s, fetched one step ahead so that its latency
  // hides behind the previous step's phases: b1 plane s+2 and np plane s+3
  // (staged in LDS), up plane s+2 at the curl positions, nu plane s+1 at the
  // E positions, b12 at the output node of plane s
  double fb[KL][3] = {}, fbase[3] = {};
  // b1 at the output node, planes s and s+1 (taken from the LDS ring)
  double b1_prev[3] = {0, 0, 0}, b1_next[3] = {0, 0, 0};
  R fn[KL] = {}, fu[KA][3] = {}, fnu[KE] = {};
  auto fetch = [&](int s) {
    const int64_t pb = wrapc(s + 2, g.nz) * nxy, pn = wrapc(s + 3, g.nz) * nxy,
                  pe = wrapc(s + 1, g.nz) * nxy;
#pragma unroll
    for (int k = 0; k < KL; k++)
      if (lp[k] >= 0) {
#pragma unroll
        for (int m = 0; m < 3; m++) fb[k][m] = b1[pb + lq[k] + m * nn];
        fn[k] = np[pn + lq[k]];
      }
#pragma unroll
    for (int k = 0; k < KA; k++)
      if (ap[k] >= 0)
#pragma unroll
        for (int m = 0; m < 3; m++) fu[k][m] = up[pb + aq[k] + m * nn];
#pragma unroll
    for (int k = 0; k < KE; k++)
      if (ep[k] >= 0) fnu[k] = nu[pe + eq[k]];
    if (oown && s >= Z0) {
      const int64_t oc = gidx(g, ox, oy, s);
#pragma unroll
      for (int m = 0; m < 3; m++) fbase[m] = b12[oc + m * nn];
    }
  };
  fetch(Z0 - 4);
  for (int s = Z0 - 4; s < Z1; s++) {
    R cu[KA][3], cnu[KE];
    double cbase[3];
    {  // stage b1 plane s+2 and np plane s+3, keep this step's up, nu, b12,
       // start the next step's loads
      const int rb = (s + 2) & 1, rn = (s + 3) & 1;
#pragma unroll
      for (int k = 0; k < KL; k++)
        if (lp[k] >= 0) {
#pragma unroll
          for (int m = 0; m < 3; m++) sb[rb][m][lp[k]] = fb[k][m];
          sn[rn][lp[k]] = fn[k];
        }
#pragma unroll
      for (int k = 0; k < KA; k++)
#pragma unroll
        for (int m = 0; m < 3; m++) cu[k][m] = fu[k][m];
#pragma unroll
      for (int k = 0; k < KE; k++) cnu[k] = fnu[k];
#pragma unroll
      for (int m = 0; m < 3; m++) cbase[m] = fbase[m];
      if (s + 1 < Z1) fetch(s + 1);
    }
    __syncthreads();
    if (s >= Z0 - 3) {  // aj, a at plane s+2 over x, y in [-1, T+1]
      const int r = (s + 2) & 1, rm = r ^ 1;
#pragma unroll
      for (int k = 0; k < KA; k++) {
        const int p = ap[k];
        if (p < 0) continue;
        const R n0 = sn[r][p];
        const R nt1 = (R)0.5 * (n0 + sn[r][p + 1]);
        const R nt2 = (R)0.5 * (n0 + sn[r][p + PW]);
        const R nt3 = (R)0.5 * (n0 + sn[rm][p]);
        const double c1 = (sb[r][2][p] - sb[r][2][p - PW]) / f.dy +
                          (sb[rm][1][p] - sb[r][1][p]) / f.dz;
        const double c2 = (sb[r][0][p] - sb[rm][0][p]) / f.dz +
                          (sb[r][2][p - 1] - sb[r][2][p]) / f.dx;
        const double c3 = (sb[r][1][p] - sb[r][1][p - 1]) / f.dx +
                          (sb[r][0][p - PW] - sb[r][0][p]) / f.dy;
        const R j1 = (R)c1 / (nt1 * al), j2 = (R)c2 / (nt2 * al), j3 = (R)c3 / (nt3 * al);
        sj[r][0][p] = j1;
        sj[r][1][p] = j2;
        sj[r][2][p] = j3;
        sa[r][0][p] = j1 - cu[k][0];
        sa[r][1][p] = j2 - cu[k][1];
        sa[r][2][p] = j3 - cu[k][2];
      }
    }
    __syncthreads();
    if (xface && s >= Z0 - 3) {
      const int r = (s + 2) & 1;
      ghost_x<PW, PA, NT>(&sj[r][0][0], 3, tid, X0, g.ncx);
      ghost_x<PW, PA, NT>(&sa[r][0][0], 3, tid, X0, g.ncx);
      __syncthreads();
    }
    if (s >= Z0 - 2) {  // ct at plane s+2 over x, y in [0, T+1]
      const int r = (s + 2) & 1, rm = r ^ 1;
      constexpr int W = TX + 2;
      for (int t = tid; t < W * (TY + 2); t += NT) {
        const int p = (t % W + 2) + PW * (t / W + 2);
        const R aa0 = (R)f.rx * (sa[r][0][p] - sa[r][0][p - 1]) + sa[r][0][p - 1];
        const R aa1 = (R)f.ry * (sa[r][1][p] - sa[r][1][p - PW]) + sa[r][1][p - PW];
        const R aa2 = (R)f.rz * (sa[r][2][p] - sa[rm][2][p]) + sa[rm][2][p];
        double hi, lo;
        hi = sb[r][0][p - PW] + f.ry * (sb[r][0][p] - sb[r][0][p - PW]);
        lo = sb[rm][0][p - PW] + f.ry * (sb[rm][0][p] - sb[rm][0][p - PW]);
        const double bx = lo + f.rz * (hi - lo);
        hi = sb[r][1][p - 1] + f.rx * (sb[r][1][p] - sb[r][1][p - 1]);
        lo = sb[rm][1][p - 1] + f.rx * (sb[rm][1][p] - sb[rm][1][p - 1]);
        const double by = lo + f.rz * (hi - lo);
        hi = sb[r][2][p - PW] + f.ry * (sb[r][2][p] - sb[r][2][p - PW]);
        lo = sb[r][2][p - 1 - PW] + f.ry * (sb[r][2][p - 1] - sb[r][2][p - 1 - PW]);
        const double bz = lo + f.rx * (hi - lo);
        const R Bx = (R)bx, By = (R)by, Bz = (R)bz;
        sc[r][0][p] = aa1 * Bz - aa2 * By;
        sc[r][1][p] = aa2 * Bx - aa0 * Bz;
        sc[r][2][p] = aa0 * By - aa1 * Bx;
      }
    }
    __syncthreads();
    if (xface && s >= Z0 - 2) {
      ghost_x<PW, PA, NT>(&sc[(s + 2) & 1][0][0], 3, tid, X0, g.ncx);
      __syncthreads();
    }
    if (s >= Z0 - 1) {  // E at plane s+1 over x, y in [0, T]
      const int r = (s + 1) & 1, rp = r ^ 1;
#pragma unroll
      for (int k = 0; k < KE; k++) {
        const int p = ep[k];
        if (p < 0) continue;
        const R nuv = cnu[k];
        se[r][0][p] = (R)0.5 * (sc[r][0][p] + sc[r][0][p + 1]) + nuv * sj[r][0][p];
        se[r][1][p] = (R)0.5 * (sc[r][1][p] + sc[r][1][p + PW]) + nuv * sj[r][1][p];
        se[r][2][p] = (R)0.5 * (sc[r][2][p] + sc[rp][2][p]) + nuv * sj[r][2][p];
      }
      // b1 of the output node at plane s+1, for the next step's output: its
      // ring slot takes plane s+3 when that step starts
      if (oown) {
        const int p = (olx + 2) + PW * (oly + 2);
#pragma unroll
        for (int m = 0; m < 3; m++) b1_next[m] = sb[r][m][p];
      }
    }
    __syncthreads();
    if (xface && s >= Z0 - 1) {
      ghost_x<PW, PA, NT>(&se[(s + 1) & 1][0][0], 3, tid, X0, g.ncx);
      __syncthreads();
    }
    if (s >= Z0 && oown) {  // b1p2 = b12 - coef curlE(E) at plane s; out = (b1p2 + b1)/2
      const int r = s & 1, rp = r ^ 1;
      const int p = (olx + 2) + PW * (oly + 2);
      const double c0 = ((double)se[r][2][p + PW] - (double)se[r][2][p]) / f.dy +
                        ((double)se[r][1][p] - (double)se[rp][1][p]) / f.dz;
      const double c1 = ((double)se[r][2][p] - (double)se[r][2][p + 1]) / f.dx +
                        ((double)se[rp][0][p] - (double)se[r][0][p]) / f.dz;
      const double c2 = ((double)se[r][0][p] - (double)se[r][0][p + PW]) / f.dy +
                        ((double)se[r][1][p + 1] - (double)se[r][1][p]) / f.dx;
      put3<double>(g, out, ox, oy, s, 0.5 * ((cbase[0] - coef * c0) + b1_prev[0]),
                   0.5 * ((cbase[1] - coef * c1) + b1_prev[1]),
                   0.5 * ((cbase[2] - coef * c2) + b1_prev[2]));
    }
#pragma unroll
    for (int m = 0; m < 3; m++) b1_prev[m] = b1_next[m];
  }
}

// the corrector's LDS rings
template <typename R, int TX, int TY>
struct CorrLds {
  static constexpr int PW = FT<TX, TY>::PW, PH = FT<TX, TY>::PH, PA = FT<TX, TY>::PA;
  double sb[2][3][PA];  // bh = (b1p2 + b1)/2, planes s+2, s+3
  R sn[2][PA];          // np, planes s+3, s+4
  R sa[3][3][PA];       // a = aj - up, planes s+1 .. s+3
  R sj[3][3][PA];       // aj
  R sw[3][3][PA];       // btc0 = edge_to_center(bh)
  R sc[2][3][PA];       // ct
  R se[2][3][PA];       // E
  int tcx[PW], tcy[PH];
};

// the corrector over tile (bx, by, bz); coef = dtsub; E and aj of the tile
// are stored when Eout is set.  DEPTH: z steps its loads run ahead.
template <typename R, int TX, int TY, int NT, int DEPTH>
__device__ __forceinline__ void sub_correct_tile(const Geo& g, const FC& f,
                                                 const double* __restrict__ bh,
                                                 const double* __restrict__ b1,
                                                 const R* __restrict__ np,
                                                 const R* __restrict__ up,
                                                 const R* __restrict__ nu, double coef, int bx,
                                                 int by, int bz, int tz,
                                                 double* __restrict__ out, R* __restrict__ Eout,
                                                 R* __restrict__ ajout, CorrLds<R, TX, TY>& L) {
  constexpr int PW = FT<TX, TY>::PW, PH = FT<TX, TY>::PH, PA = FT<TX, TY>::PA;
  static_assert(NT >= TX * TY, "one thread per output column at least");
  auto& sb = L.sb;
  auto& sn = L.sn;
  auto& sa = L.sa;
  auto& sj = L.sj;
  auto& sw = L.sw;
  auto& sc = L.sc;
  auto& se = L.se;
  auto& tcx = L.tcx;
  auto& tcy = L.tcy;
  const int tid = threadIdx.x;
  const int X0 = 1 + bx * TX, Y0 = 1 + by * TY, Z0 = 1 + bz * tz;
  const int Z1 = min(Z0 + tz, g.ncz + 1);
  if (Z0 > g.ncz) return;  // empty z chunk (whole workgroup)
  for (int t = tid; t < PW; t += NT) tcx[t] = tile_cx(g, X0 - 2 + t);
  for (int t = tid; t < PH; t += NT) tcy[t] = wrapc(Y0 - 2 + t, g.ny);
  __syncthreads();
  const int64_t nxy = (int64_t)g.nx * g.ny, nn = g.nn;
  const R al = (R)f.alpha;
  // an open x and a tile whose halo reaches an x ghost (uniform per workgroup)
  const bool xface = g.ox && (X0 - 2 < 1 || X0 + TX + 2 > g.ncx);
  constexpr int AW = TX + 3, AN = AW * (TY + 3), EW = TX + 1, EN = EW * (TY + 1);
  constexpr int KL = (PA + NT - 1) / NT, KA = (AN + NT - 1) / NT, KE = (EN + NT - 1) / NT;
  int lp[KL], lq[KL], ap[KA], aq[KA], ep[KE], eq[KE];
  own<KL, NT, PW>(tid, PW, PA, 0, 0, tcx, tcy, g.nx, lp, lq);
  own<KA, NT, PW>(tid, AW, AN, 1, 1, tcx, tcy, g.nx, ap, aq);
  own<KE, NT, PW>(tid, EW, EN, 2, 2, tcx, tcy, g.nx, ep, eq);
  const int olx = tid % TX, oly = tid / TX, ox = X0 + olx, oy = Y0 + oly;
  const bool oown = tid < TX * TY && ox <= g.ncx && oy <= g.ncy;
  // step s's global data, fetched DEPTH steps ahead into DEPTH register
  // sets that the z loop takes in turn: bh plane s+3 and np plane s+4
  // (staged), up plane s+3 at the curl positions, nu plane s+1 at the E
  // positions, b1 at the output node of plane s
  struct Fetched {
    double fp[KL][3], fbase[3];
    R fn[KL], fu[KA][3], fnu[KE];
  } fs[DEPTH]{};
  auto fetch = [&](Fetched& F, int s) {
    const int64_t pb = wrapc(s + 3, g.nz) * nxy, pn = wrapc(s + 4, g.nz) * nxy,
                  pe = wrapc(s + 1, g.nz) * nxy;
#pragma unroll
    for (int k = 0; k < KL; k++)
      if (lp[k] >= 0) {
#pragma unroll
        for (int m = 0; m < 3; m++) F.fp[k][m] = bh[pb + lq[k] + m * nn];
        F.fn[k] = np[pn + lq[k]];
      }
#pragma unroll
    for (int k = 0; k < KA; k++)
      if (ap[k] >= 0)
#pragma unroll
        for (int m = 0; m < 3; m++) F.fu[k][m] = up[pb + aq[k] + m * nn];
#pragma unroll
    for (int k = 0; k < KE; k++)
      if (ep[k] >= 0) F.fnu[k] = nu[pe + eq[k]];
    if (oown && s >= Z0) {
      const int64_t oc = gidx(g, ox, oy, s);
#pragma unroll
      for (int m = 0; m < 3; m++) F.fbase[m] = b1[oc + m * nn];
    }
  };
#pragma unroll
  for (int d = 0; d < DEPTH; d++)
    if (Z0 - 5 + d < Z1) fetch(fs[d], Z0 - 5 + d);
  auto step = [&](const int s, Fetched& F) {
    R cu[KA][3], cnu[KE];
    double cbase[3];
    {  // stage bh plane s+3 and np plane s+4, keep this step's up, nu, b1,
       // start the next step's loads
      const int rb = (s + 3) & 1, rn = (s + 4) & 1;
#pragma unroll
      for (int k = 0; k < KL; k++)
        if (lp[k] >= 0) {
#pragma unroll
          for (int m = 0; m < 3; m++) sb[rb][m][lp[k]] = F.fp[k][m];
          sn[rn][lp[k]] = F.fn[k];
        }
#pragma unroll
      for (int k = 0; k < KA; k++)
#pragma unroll
        for (int m = 0; m < 3; m++) cu[k][m] = F.fu[k][m];
#pragma unroll
      for (int k = 0; k < KE; k++) cnu[k] = F.fnu[k];
#pragma unroll
      for (int m = 0; m < 3; m++) cbase[m] = F.fbase[m];
      if (s + DEPTH < Z1) fetch(F, s + DEPTH);
    }
    __syncthreads();
    if (s >= Z0 - 4) {  // aj, a over [-1, T+1] and btc0 over [-1, T+2], plane s+3
      const int r = (s + 3) & 1, rm = r ^ 1, w = r3(s + 3);
#pragma unroll
      for (int k = 0; k < KA; k++) {
        const int p = ap[k];
        if (p < 0) continue;
        const R n0 = sn[r][p];
        const R nt1 = (R)0.5 * (n0 + sn[r][p + 1]);
        const R nt2 = (R)0.5 * (n0 + sn[r][p + PW]);
        const R nt3 = (R)0.5 * (n0 + sn[rm][p]);
        const double c1 = (sb[r][2][p] - sb[r][2][p - PW]) / f.dy +
                          (sb[rm][1][p] - sb[r][1][p]) / f.dz;
        const double c2 = (sb[r][0][p] - sb[rm][0][p]) / f.dz +
                          (sb[r][2][p - 1] - sb[r][2][p]) / f.dx;
        const double c3 = (sb[r][1][p] - sb[r][1][p - 1]) / f.dx +
                          (sb[r][0][p - PW] - sb[r][0][p]) / f.dy;
        const R j1 = (R)c1 / (nt1 * al), j2 = (R)c2 / (nt2 * al), j3 = (R)c3 / (nt3 * al);
        sj[w][0][p] = j1;
        sj[w][1][p] = j2;
        sj[w][2][p] = j3;
        sa[w][0][p] = j1 - cu[k][0];
        sa[w][1][p] = j2 - cu[k][1];
        sa[w][2][p] = j3 - cu[k][2];
      }
      constexpr int W2 = TX + 4;
      for (int t = tid; t < W2 * (TY + 4); t += NT) {
        const int p = (t % W2 + 1) + PW * (t / W2 + 1);
        double hi, lo;
        hi = sb[r][0][p - PW] + f.ry * (sb[r][0][p] - sb[r][0][p - PW]);
        lo = sb[rm][0][p - PW] + f.ry * (sb[rm][0][p] - sb[rm][0][p - PW]);
        sw[w][0][p] = (R)(lo + f.rz * (hi - lo));
        hi = sb[r][1][p - 1] + f.rx * (sb[r][1][p] - sb[r][1][p - 1]);
        lo = sb[rm][1][p - 1] + f.rx * (sb[rm][1][p] - sb[rm][1][p - 1]);
        sw[w][1][p] = (R)(lo + f.rz * (hi - lo));
        hi = sb[r][2][p - PW] + f.ry * (sb[r][2][p] - sb[r][2][p - PW]);
        lo = sb[r][2][p - 1 - PW] + f.ry * (sb[r][2][p - 1] - sb[r][2][p - 1 - PW]);
        sw[w][2][p] = (R)(lo + f.rx * (hi - lo));
      }
    }
    __syncthreads();
    if (xface && s >= Z0 - 4) {
      const int w = r3(s + 3);
      ghost_x<PW, PA, NT>(&sj[w][0][0], 3, tid, X0, g.ncx);
      ghost_x<PW, PA, NT>(&sa[w][0][0], 3, tid, X0, g.ncx);
      ghost_x<PW, PA, NT>(&sw[w][0][0], 3, tid, X0, g.ncx);
      __syncthreads();
    }
    if (s >= Z0 - 2) {  // ct at plane s+2 over [0, T+1]: a x smoothed B
      const int r = (s + 2) & 1, wc = r3(s + 2), wm = r3(s + 1), wp = r3(s + 3);
      constexpr int W = TX + 2;
      for (int t = tid; t < W * (TY + 2); t += NT) {
        const int p = (t % W + 2) + PW * (t / W + 2);
        const R aa0 = (R)f.rx * (sa[wc][0][p] - sa[wc][0][p - 1]) + sa[wc][0][p - 1];
        const R aa1 = (R)f.ry * (sa[wc][1][p] - sa[wc][1][p - PW]) + sa[wc][1][p - PW];
        const R aa2 = (R)f.rz * (sa[wc][2][p] - sa[wm][2][p]) + sa[wm][2][p];
        const R m1c = (R)0.5 * (sw[wc][0][p] + sw[wc][0][p + 1]);
        // on an open x the face average below node 1 is a copy of node 1's
        const R m1m = g.ox && X0 - 2 + p % PW == 1 ? m1c
                                                   : (R)0.5 * (sw[wc][0][p - 1] + sw[wc][0][p]);
        const R m2c = (R)0.5 * (sw[wc][1][p] + sw[wc][1][p + PW]);
        const R m2m = (R)0.5 * (sw[wc][1][p - PW] + sw[wc][1][p]);
        const R m3c = (R)0.5 * (sw[wc][2][p] + sw[wp][2][p]);
        const R m3m = (R)0.5 * (sw[wm][2][p] + sw[wc][2][p]);
        const R bx = (R)f.rx * (m1c - m1m) + m1m;
        const R by = (R)f.ry * (m2c - m2m) + m2m;
        const R bz = (R)f.rz * (m3c - m3m) + m3m;
        sc[r][0][p] = aa1 * bz - aa2 * by;
        sc[r][1][p] = aa2 * bx - aa0 * bz;
        sc[r][2][p] = aa0 * by - aa1 * bx;
      }
    }
    __syncthreads();
    if (xface && s >= Z0 - 2) {
      ghost_x<PW, PA, NT>(&sc[(s + 2) & 1][0][0], 3, tid, X0, g.ncx);
      __syncthreads();
    }
    if (s >= Z0 - 1) {  // E at plane s+1 over [0, T]
      const int r = (s + 1) & 1, rp = r ^ 1, w = r3(s + 1);
      const int ze = wrapc(s + 1, g.nz);
      const bool store = Eout && s + 1 >= Z0 && s + 1 < Z1;
#pragma unroll
      for (int k = 0; k < KE; k++) {
        const int p = ep[k];
        if (p < 0) continue;
        const int lx = p % PW - 2, ly = p / PW - 2;
        const R nuv = cnu[k];
        const R e0 = (R)0.5 * (sc[r][0][p] + sc[r][0][p + 1]) + nuv * sj[w][0][p];
        const R e1 = (R)0.5 * (sc[r][1][p] + sc[r][1][p + PW]) + nuv * sj[w][1][p];
        const R e2 = (R)0.5 * (sc[r][2][p] + sc[rp][2][p]) + nuv * sj[w][2][p];
        se[r][0][p] = e0;
        se[r][1][p] = e1;
        se[r][2][p] = e2;
        if (store && lx < TX && ly < TY && X0 + lx <= g.ncx && Y0 + ly <= g.ncy) {
          put3<R>(g, Eout, X0 + lx, Y0 + ly, ze, e0, e1, e2);
          put3<R>(g, ajout, X0 + lx, Y0 + ly, ze, sj[w][0][p], sj[w][1][p], sj[w][2][p]);
        }
      }
    }
    __syncthreads();
    if (xface && s >= Z0 - 1) {
      ghost_x<PW, PA, NT>(&se[(s + 1) & 1][0][0], 3, tid, X0, g.ncx);
      __syncthreads();
    }
    if (s >= Z0 && oown) {  // b = b1 - coef curlE(E) at plane s
      const int r = s & 1, rp = r ^ 1;
      const int p = (olx + 2) + PW * (oly + 2);
      const double c0 = ((double)se[r][2][p + PW] - (double)se[r][2][p]) / f.dy +
                        ((double)se[r][1][p] - (double)se[rp][1][p]) / f.dz;
      const double c1 = ((double)se[r][2][p] - (double)se[r][2][p + 1]) / f.dx +
                        ((double)se[rp][0][p] - (double)se[r][0][p]) / f.dz;
      const double c2 = ((double)se[r][0][p] - (double)se[r][0][p + PW]) / f.dy +
                        ((double)se[r][1][p + 1] - (double)se[r][1][p]) / f.dx;
      put3<double>(g, out, ox, oy, s, cbase[0] - coef * c0, cbase[1] - coef * c1,
                   cbase[2] - coef * c2);
    }
  };
  for (int s = Z0 - 5; s < Z1; s += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++)
      if (s + d < Z1) step(s + d, fs[d]);
  }
}

// z steps that the corrector's loads run ahead of their use: the corrector
// holds one workgroup per CU (its LDS), 2 waves per SIMD in FP32, so the
// registers are there for DEPTH fetched planes, and its z march was bound
// by the latency of its loads (the field stage at 256x128x128, 10
// subcycles: one step ahead 3.90 ms, two 3.60, three 3.55, four 3.49)
template <typename R>
struct CorrDepth {
  static constexpr int value = 4;
};

// one predictor launch: a workgroup per tile; coef = cmul dtsub
template <typename R, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_sub_predict(Geo g, FC f, const double* __restrict__ b1,
                                                        const double* __restrict__ b12,
                                                        const R* __restrict__ np,
                                                        const R* __restrict__ up,
                                                        const R* __restrict__ nu,
                                                        const double* __restrict__ ts, double cmul,
                                                        int tz, double* __restrict__ out) {
  __shared__ PredLds<R, TX, TY> L;
  sub_predict_tile<R, TX, TY, NT>(g, f, b1, b12, np, up, nu, cmul * ts[0], blockIdx.x,
                                  blockIdx.y, blockIdx.z, tz, out, L);
}

// one corrector launch
template <typename R, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_sub_correct(Geo g, FC f, const double* __restrict__ bh,
                                                        const double* __restrict__ b1,
                                                        const R* __restrict__ np,
                                                        const R* __restrict__ up,
                                                        const R* __restrict__ nu,
                                                        const double* __restrict__ ts, double cmul,
                                                        int tz, double* __restrict__ out,
                                                        R* __restrict__ Eout,
                                                        R* __restrict__ ajout) {
  __shared__ CorrLds<R, TX, TY> L;
  sub_correct_tile<R, TX, TY, NT, CorrDepth<R>::value>(g, f, bh, b1, np, up, nu, cmul * ts[0], blockIdx.x,
                                  blockIdx.y, blockIdx.z, tz, out, Eout, ajout, L);
}

// nsub subcycles in one cooperative launch of as many workgroups as the chip
// holds at once: the workgroups take the predictor's tiles in turn, meet at a
// grid barrier (whose fences write back and invalidate every XCD's L2, so
// each phase reads what the last one wrote on any XCD), take the
// corrector's tiles, meet again, and so on.  The B buffers rotate as the
// host rotates them after a two-launch subcycle: b1 and b12 trade places,
// bh is the predictor's output every time.  Tiles are numbered x fastest,
// then y, then z, as the two-launch grid dispatches them.
template <typename R, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_subcycles(Geo g, FC f, double* b1, double* b12,
                                                      double* bh, const R* np, const R* up,
                                                      const R* nu, const double* ts, int nsub,
                                                      int tz, R* Eout, R* ajout) {
  __shared__ union {
    PredLds<R, TX, TY> p;
    CorrLds<R, TX, TY> c;
  } L;
  const int gx = (g.ncx + TX - 1) / TX, gy = (g.ncy + TY - 1) / TY;
  const int nt = gx * gy * ((g.ncz + tz - 1) / tz);
  const double dts = ts[0];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  for (int q = 0; q < nsub; q++) {
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
      __syncthreads();  // the last tile's reads of the rings are done
      sub_predict_tile<R, TX, TY, NT>(g, f, b1, b12, np, up, nu, 2.0 * dts, t % gx,
                                      (t / gx) % gy, t / (gx * gy), tz, bh, L.p);
    }
    grid.sync();
    const bool last = q == nsub - 1;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
      __syncthreads();
      sub_correct_tile<R, TX, TY, NT, CorrDepth<R>::value>(g, f, bh, b1, np, up, nu, 1.0 * dts, t % gx,
                                      (t / gx) % gy, t / (gx * gy), tz, b12,
                                      last ? Eout : nullptr, last ? ajout : nullptr, L.c);
    }
    if (!last) grid.sync();
    double* const t = b1;
    b1 = b12;
    b12 = t;
  }
}

template <typename R>
R* F(swap_ctx* c, int id) {
  return (R*)c->fld[id];
}

}  // namespace

#define DISPATCH(c, ...)                        \
  do {                                          \
    if ((c)->prec == 64) {                      \
      using R = double;                         \
      __VA_ARGS__;                              \
    } else {                                    \
      using R = float;                          \
      __VA_ARGS__;                              \
    }                                           \
  } while (0)

#define LAUNCH(kern, ...) hipLaunchKernelGGL(kern, ggrid(c->g), gblock(), 0, c->stream, __VA_ARGS__)

static int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

int fl_curlB(swap_ctx* c, const double* b, const double* b2, void* aj, void* a) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_curlB<R>), c->g, f, b, b2, F<R>(c, SWAP_F_NP), F<R>(c, SWAP_F_UP), (R*)aj,
                     (R*)a));
  return post(c);
}
int fl_cross_nodes(swap_ctx* c, const void* a, const double* b, void* ct) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_cross_nodes<R>), c->g, f, (const R*)a, b, (R*)ct));
  return post(c);
}
int fl_e2c_B(swap_ctx* c, const double* b, const double* b2, void* out) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_e2c_B<R>), c->g, f, b, b2, (R*)out));
  return post(c);
}
int fl_cross_smooth(swap_ctx* c, const void* a, const void* btc0, void* ct) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_cross_smooth<R>), c->g, f, (const R*)a, (const R*)btc0, (R*)ct));
  return post(c);
}
int fl_edgeavg_E(swap_ctx* c, const void* ct, const void* aj, void* E) {
  DISPATCH(c, LAUNCH((k_edgeavg_E<R>), c->g, (const R*)ct, (const R*)aj, F<R>(c, SWAP_F_NU),
                     (R*)E));
  return post(c);
}
int fl_curlE_update(swap_ctx* c, const void* E, const double* base, double coef, double* out) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_curlE_update<R>), c->g, f, (const R*)E, base, coef, out));
  return post(c);
}
int fl_curlE(swap_ctx* c, const void* E, void* out) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_curlE<R>), c->g, f, (const R*)E, (R*)out));
  return post(c);
}
int fl_f2c(swap_ctx* c, const void* in, void* out) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_f2c<R>), c->g, f, (const R*)in, (R*)out));
  return post(c);
}
int fl_e2c(swap_ctx* c, const void* in, void* out) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_e2c<R>), c->g, f, (const R*)in, (R*)out));
  return post(c);
}
int fl_e2f(swap_ctx* c, const void* in, const double* bin, void* out) {
  FC f = fc_of(c);
  DISPATCH(c, if (bin) LAUNCH((k_e2c_B<R>), c->g, f, bin, (const double*)nullptr, (R*)c->btc0);
           else LAUNCH((k_e2c<R>), c->g, f, (const R*)in, (R*)c->btc0);
           LAUNCH((k_face_avg<R>), c->g, (const R*)c->btc0, (R*)out));
  return post(c);
}
int fl_crossf2(swap_ctx* c, const void* aa, const void* btc, void* out) {
  DISPATCH(c, LAUNCH((k_crossf2<R>), c->g, (const R*)aa, (const R*)btc, (R*)out));
  return post(c);
}
int fl_periodic(swap_ctx* c, void* f, int ncomp, int is_double) {
  if (is_double || c->prec == 64)
    LAUNCH((k_periodic<double>), c->g, (double*)f, ncomp);
  else
    LAUNCH((k_periodic<float>), c->g, (float*)f, ncomp);
  return post(c);
}
int fl_a_from(swap_ctx* c, const void* aj, const void* up, void* a) {
  DISPATCH(c, LAUNCH((k_a_from<R>), c->g, (const R*)aj, (const R*)up, (R*)a));
  return post(c);
}
int fl_build_src(swap_ctx* c, const void* aj, const void* up, const void* btc, void* ajc,
                 void* S) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_build_src<R>), c->g, f, (const R*)aj, (const R*)up, (const R*)btc,
                     (R*)ajc, F<R>(c, SWAP_F_UPC), (R*)S));
  return post(c);
}
int fl_update_src(swap_ctx* c, const void* up, const void* ajc, void* S) {
  FC f = fc_of(c);
  DISPATCH(c, LAUNCH((k_update_src<R>), c->g, f, (const R*)up, (const R*)ajc,
                     F<R>(c, SWAP_F_UPC), (R*)S));
  return post(c);
}
int fl_finalize_np(swap_ctx* c) {
  DISPATCH(c, using A = typename AccOf<R>::T;
           LAUNCH((k_finalize_np<R>), c->g, (const A*)c->macc, (R)c->qn, F<R>(c, SWAP_F_NP)));
  return post(c);
}
int fl_finalize_up(swap_ctx* c) {
  DISPATCH(c, using A = typename AccOf<R>::T;
           LAUNCH((k_finalize_up<R>), c->g, (const A*)c->macc + c->g.nn, F<R>(c, SWAP_F_NP),
                  (R)c->qn, F<R>(c, SWAP_F_UP)));
  return post(c);
}
// check_time_step (gutsf.f:1776-1822) on the device: per-node deltat and
// its grid minimum (k_deltat), then dtsub and ntf into c->ts (k_tstep), with
// maind.f:420-421's reset dtsub = dt/ntsub, ntf = ntsub first
int fl_time_step_dev(swap_ctx* c) {
  FC f = fc_of(c);
  const double dtsub0 = c->cfg.dt / (double)c->cfg.ntsub;
  const float big = 3.0e38f;
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(1), 0, c->stream, c->dmin, big);
  const unsigned nb = (unsigned)std::min<int64_t>(DT_BLOCKS, (c->g.nn + 255) / 256);
  DISPATCH(c, hipLaunchKernelGGL((k_deltat<R>), dim3(nb), dim3(256), 0, c->stream, c->g, f,
                                 (const double*)c->b1, (const R*)F<R>(c, SWAP_F_NP),
                                 (R*)c->scratch, c->dmin));
  // every node passes unless some deltat <= 2 dtsub0 (compared in the
  // context precision, as the per-node test is)
  const double thr64 = 2.0 * dtsub0 * (1 + 1e-6);
  const float thr32 = 2.0f * (float)dtsub0 * (1 + 1e-6f);
  DISPATCH(c, hipLaunchKernelGGL((k_tstep<R>), dim3(1), dim3(1024), 0, c->stream, c->g,
                                 (const R*)c->scratch, (const float*)c->dmin, dtsub0, thr64,
                                 thr32, c->cfg.ntsub, c->ts));
  return post(c);
}
// the same, handed back to the host (swap_check_time_step)
int fl_check_time_step(swap_ctx* c, int* ntf, double* dtsub) {
  if (int e = fl_time_step_dev(c)) return e;
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, c->ts, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&c->box_over, c->dcnt + 4, sizeof(c->box_over),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *dtsub = h[0];
  *ntf = (int)h[1];
  return 0;
}
// dtsub of swap_field_subcycles' caller onto the device
int fl_set_dtsub(swap_ctx* c, double dtsub, int ntf) {
  hipLaunchKernelGGL(k_set_dtsub, dim3(1), dim3(1), 0, c->stream, c->ts, dtsub, (double)ntf);
  return post(c);
}
int fl_copy_b(swap_ctx* c, const double* src, double* dst) {
  HIPCHK(c, hipMemcpyAsync(dst, src, sizeof(double) * 3 * c->g.nn, hipMemcpyDeviceToDevice,
                           c->stream));
  return 0;
}
// one fused subcycle: b1p2 = predictor(b1, b12); b12's buffer = corrector
// (b1p2, b1); E and aj of the corrector are stored when last is set
// tile of canonical columns per workgroup and threads of the predictor (NP)
// and the corrector (NC).  FP32: 32x8 columns and 512 threads measured best
// among 16..64 x 4..12 columns and 256..1024 threads at 256x128x128 cells
// (more threads than columns hide the latency of each z step's phases).  The
// z chunks follow from occupancy (z_chunks): there the predictor holds two
// workgroups per CU and marches 32 planes, the corrector (more LDS) one and
// 64 planes, 0.11 ms/step less than 32 planes for both.  The FP64 build's
// LDS rings take twice the bytes per position.
template <typename R>
struct FTile {
  static constexpr int TX = 32, TY = 8, NP = 512, NC = 512;
};
template <>
struct FTile<double> {
  static constexpr int TX = 32, TY = 4, NP = 128, NC = 128;
};

// z chunks of a fused launch: as many workgroups as the chip holds at once
// (CUs x the kernel's resident workgroups per CU, which its LDS sets), so a
// launch runs in one round: with each workgroup marching its z chunk, a
// second round of workgroups would double the time of a launch whose chip
// was full in the first.  Computed once per context and kernel.
template <typename Kern>
int z_chunks(swap_ctx* c, Kern kern, int nthreads, int nxy, int* cache) {
  if (*cache > 0) return *cache;
  int per = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, nthreads, 0) != hipSuccess ||
      per < 1)
    per = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) !=
          hipSuccess ||
      cus < 1)
    cus = 256;
  *cache = std::max(1, std::min((cus * per) / std::max(nxy, 1), c->g.ncz));
  return *cache;
}

int fl_subcycle_fused(swap_ctx* c, bool last) {
  FC f = fc_of(c);
  const Geo& g = c->g;
  DISPATCH(c, constexpr int TX = FTile<R>::TX, TY = FTile<R>::TY;
           constexpr int NP = FTile<R>::NP, NC = FTile<R>::NC;
           const int gx = (g.ncx + TX - 1) / TX, gy = (g.ncy + TY - 1) / TY;
           const int zp = z_chunks(c, k_sub_predict<R, TX, TY, NP>, NP, gx * gy, &c->fz_pred);
           const int zc = z_chunks(c, k_sub_correct<R, TX, TY, NC>, NC, gx * gy, &c->fz_corr);
           const int tzp = (g.ncz + zp - 1) / zp, tzc = (g.ncz + zc - 1) / zc;
           hipLaunchKernelGGL((k_sub_predict<R, TX, TY, NP>), dim3(gx, gy, (g.ncz + tzp - 1) / tzp),
                              dim3(NP), 0, c->stream, g, f,
                                 (const double*)c->b1, (const double*)c->b12,
                                 (const R*)F<R>(c, SWAP_F_NP), (const R*)F<R>(c, SWAP_F_UP),
                                 (const R*)F<R>(c, SWAP_F_NU), (const double*)c->ts, 2.0, tzp,
                                 c->b1p2);
           hipLaunchKernelGGL((k_sub_correct<R, TX, TY, NC>), dim3(gx, gy, (g.ncz + tzc - 1) / tzc),
                              dim3(NC), 0, c->stream, g, f,
                              (const double*)c->b1p2, (const double*)c->b1,
                              (const R*)F<R>(c, SWAP_F_NP), (const R*)F<R>(c, SWAP_F_UP),
                              (const R*)F<R>(c, SWAP_F_NU), (const double*)c->ts, 1.0, tzc,
                              c->b12,
                              last ? F<R>(c, SWAP_F_E) : (R*)nullptr,
                              last ? F<R>(c, SWAP_F_AJ) : (R*)nullptr));
  return post(c);
}
// nsub fused subcycles in one cooperative launch (k_subcycles): one
// workgroup per CU at the corrector's LDS, z chunks so that the tiles of a
// phase match the workgroups; E and aj are stored by the last corrector.
// The caller rotates the B buffers as after nsub two-launch subcycles.
int fl_subcycles_persistent(swap_ctx* c, int nsub) {
  if (nsub < 1) return 0;
  FC f = fc_of(c);
  Geo g = c->g;
  DISPATCH(c, constexpr int TX = FTile<R>::TX, TY = FTile<R>::TY, NT = FTile<R>::NC;
           auto kern = k_subcycles<R, TX, TY, NT>;
           const int gx = (g.ncx + TX - 1) / TX, gy = (g.ncy + TY - 1) / TY;
           if (c->pers_grid <= 0) {
             int per = 0, cus = 0;
             HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, NT, 0));
             HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                             c->device));
             if (per < 1 || cus < 1)
               return set_err(c, SWAP_EHIP, "k_subcycles does not fit a CU");
             const int nz = std::max(1, std::min((cus * per) / (gx * gy), g.ncz));
             c->pers_tz = (g.ncz + nz - 1) / nz;
             const int nt = gx * gy * ((g.ncz + c->pers_tz - 1) / c->pers_tz);
             c->pers_grid = std::min(cus * per, nt);
           }
           double* b1 = c->b1;
           double* b12 = c->b12;
           double* bh = c->b1p2;
           const R* np = F<R>(c, SWAP_F_NP);
           const R* up = F<R>(c, SWAP_F_UP);
           const R* nu = F<R>(c, SWAP_F_NU);
           const double* ts = c->ts;
           int tz = c->pers_tz;
           R* E = F<R>(c, SWAP_F_E);
           R* aj = F<R>(c, SWAP_F_AJ);
           void* args[] = {&g, &f, &b1, &b12, &bh, &np, &up, &nu, &ts, &nsub, &tz, &E, &aj};
           HIPCHK(c, hipLaunchCooperativeKernel((const void*)kern, dim3(c->pers_grid), dim3(NT),
                                                args, 0, c->stream)));
  return post(c);
}
int fl_zero(swap_ctx* c, void* p, size_t bytes) {
  HIPCHK(c, hipMemsetAsync(p, 0, bytes, c->stream));
  return 0;
}
