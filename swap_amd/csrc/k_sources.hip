// k_sources.hip -- particle loading, pickup-ion sources, FP64 diagnostics.
#include <algorithm>
#include <cmath>
#include "particle_math.h"

namespace {

constexpr double PI_REF = 3.14159;  // the reference's PI (inputs.f:52)

// absolute position x (km, uniform grid with q(i) = i d) -> cell + fraction,
// right-closed as get_interp_weights' search (gutsp.f:1394-1412)
__device__ __forceinline__ void split_pos(double x, double d, int nc, int& cell, double& f) {
  double t = x / d;          // q(i) = i d
  double i = ceil(t) - 1.0;  // Fortran cell index i: q(i) < x <= q(i+1)
  f = t - i;
  int c = (int)i - 1;        // 0-based cell
  c %= nc;
  if (c < 0) c += nc;
  cell = c;
}

template <typename R>
__global__ void k_load_abs(int64_t n, int64_t off, PDev<R> P, const double* xp, const double* vp,
                           const double* vp1, double dx, double dy, double dz, int ncx, int ncy,
                           int ncz) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  int ci, cj, ck;
  double f0, f1, f2;
  split_pos(xp[p], dx, ncx, ci, f0);
  split_pos(xp[p + n], dy, ncy, cj, f1);
  split_pos(xp[p + 2 * n], dz, ncz, ck, f2);
  int64_t q = off + p;
  P.cell[q] = pack_cell(ci, cj, ck);
  P.fx[q] = (R)f0; P.fy[q] = (R)f1; P.fz[q] = (R)f2;
  P.vx[q] = (R)vp[p]; P.vy[q] = (R)vp[p + n]; P.vz[q] = (R)vp[p + 2 * n];
  P.v1x[q] = (R)vp1[p]; P.v1y[q] = (R)vp1[p + n]; P.v1z[q] = (R)vp1[p + 2 * n];
}

template <typename R>
__global__ void k_set_abs(int64_t n, PDev<R> P, const double* xp, double dx, double dy, double dz,
                          int ncx, int ncy, int ncz) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  int ci, cj, ck;
  double f0, f1, f2;
  split_pos(xp[p], dx, ncx, ci, f0);
  split_pos(xp[p + n], dy, ncy, cj, f1);
  split_pos(xp[p + 2 * n], dz, ncz, ck, f2);
  P.cell[p] = pack_cell(ci, cj, ck);
  P.fx[p] = (R)f0; P.fy[p] = (R)f1; P.fz[p] = (R)f2;
}

template <typename R>
__global__ void k_export_abs(int64_t n, PDev<R> P, double* xp, double dx, double dy, double dz) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  xp[p] = ((double)cell_i(c) + 1.0 + (double)P.fx[p]) * dx;
  xp[p + n] = ((double)cell_j(c) + 1.0 + (double)P.fy[p]) * dy;
  xp[p + 2 * n] = ((double)cell_k(c) + 1.0 + (double)P.fz[p]) * dz;
}

// load_Maxwellian's sampler (part_init.f:176-206): position uniform on
// (q(1), q(n-1)] from 1-u, velocity vsw + vth sqrt(-ln u1) cos(PI u2) per
// component, vp1 = vp; uniforms from Philox-4x32-10 keyed by the seed, with
// counter (id lo, id hi, stream, draw)
template <typename R>
__global__ void k_maxwellian(int64_t n, int64_t off, PDev<R> P, uint32_t k0, uint32_t k1,
                             uint32_t stream, double vsw, double vth, int ncx, int ncy, int ncz,
                             float bpv, float mrv, float* bp, float* mr, uint32_t* id) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint64_t gid = (uint64_t)(off + p);
  uint32_t r[3][4];
  for (int d = 0; d < 3; d++) {
    r[d][0] = (uint32_t)gid;
    r[d][1] = (uint32_t)(gid >> 32);
    r[d][2] = stream;
    r[d][3] = (uint32_t)d;
    philox4x32_10(r[d], k0, k1);
  }
  int nc[3] = {ncx, ncy, ncz};
  int ijk[3];
  double f[3];
  for (int a = 0; a < 3; a++) {
    double t = (1.0 - u24(r[0][a])) * (double)nc[a];  // exact in double
    double i = ceil(t) - 1.0;
    ijk[a] = (int)i;
    f[a] = t - i;
  }
  double v[3];
  const uint32_t* ur[6] = {&r[0][3], &r[1][0], &r[1][1], &r[1][2], &r[1][3], &r[2][0]};
  for (int a = 0; a < 3; a++) {
    double u1 = 1.0 - u24(*ur[2 * a]);  // (0,1]
    double u2 = u24(*ur[2 * a + 1]);
    v[a] = (a == 0 ? vsw : 0.0) + vth * sqrt(-log(u1)) * cos(PI_REF * u2);
  }
  int64_t q = off + p;
  P.cell[q] = pack_cell(ijk[0], ijk[1], ijk[2]);
  P.fx[q] = (R)f[0]; P.fy[q] = (R)f[1]; P.fz[q] = (R)f[2];
  P.vx[q] = (R)v[0]; P.vy[q] = (R)v[1]; P.vz[q] = (R)v[2];
  P.v1x[q] = (R)v[0]; P.v1y[q] = (R)v[1]; P.v1z[q] = (R)v[2];
  if (bp) bp[q] = bpv;
  if (mr) mr[q] = mrv;
  if (id) id[q] = (uint32_t)gid;
}

// Ionize_sw_mp's new ions (chem_rates.f:567-608): ring velocity
// (vsw + vsw cos th, vsw sin th, 0) with th = 2 PI u, uniform position,
// mrat = 1, beta_p = beta_pu
template <typename R>
__global__ void k_ionize_sw(int64_t n, int64_t off, PDev<R> P, uint32_t k0, uint32_t k1,
                            uint32_t stream, double vsw, int ncx, int ncy, int ncz, float beta_pu,
                            float* bp, float* mr, uint32_t* id, uint32_t id0) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t r[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 7u};
  philox4x32_10(r, k0, k1);
  double th = u24(r[0]) * 2.0 * PI_REF;
  int nc[3] = {ncx, ncy, ncz};
  int ijk[3];
  double f[3];
  for (int a = 0; a < 3; a++) {
    double t = (1.0 - u24(r[1 + a])) * (double)nc[a];
    double i = ceil(t) - 1.0;
    ijk[a] = (int)i;
    f[a] = t - i;
  }
  double v[3] = {vsw + vsw * cos(th), vsw * sin(th), 0.0};
  int64_t q = off + p;
  P.cell[q] = pack_cell(ijk[0], ijk[1], ijk[2]);
  P.fx[q] = (R)f[0]; P.fy[q] = (R)f[1]; P.fz[q] = (R)f[2];
  P.vx[q] = (R)v[0]; P.vy[q] = (R)v[1]; P.vz[q] = (R)v[2];
  P.v1x[q] = (R)v[0]; P.v1y[q] = (R)v[1]; P.v1z[q] = (R)v[2];
  if (bp) bp[q] = beta_pu;
  if (mr) mr[q] = 1.0f;
  if (id) id[q] = id0 + (uint32_t)p;
}

// Inflow through the open x face at cell coordinate 0 (build-defined; the
// reference instead refills its boundary layer to the upstream density,
// check_min_den_boundary, gutsp.f:575-816).  Part of swap_step, right after
// the particle pass: the ions of an upstream drifting Maxwellian that cross
// the face during the step.  Slot s of the slab (-L, 0] x (0, ncy] x (0, ncz]
// (cell units, positions at the start of the step) draws a position and a
// velocity as the loader does and makes the pass's two half moves: an ion
// inside after the first deposits there with the pass's moments, one inside
// after the second is appended.  Philox counter (slot, step, stream, draw);
// explicit fma so that the host restatement rounds identically.
struct InflowDev {
  double vsw, vth, sh[3];  // sh: dt / 2 / cell size per axis
  int L, nc[3];
  uint32_t spc;  // slots per y-z column of the face
};
struct InflowIon {
  double th[3], te[3], v[3];  // cell coordinates after each half move
  bool dep, keep;
};
__device__ __forceinline__ double wrap_open(double t, double nc) {
  if (t > nc || t <= 0.0) {  // back into (0, nc]
    t -= (ceil(t / nc) - 1.0) * nc;
    if (t <= 0.0) t += nc;
    else if (t > nc) t -= nc;
  }
  return t;
}
__device__ __forceinline__ InflowIon inflow_slot(const InflowDev& d, uint32_t slot, uint32_t step,
                                                 uint32_t k0, uint32_t k1, uint32_t stream) {
#pragma clang fp contract(off)
  uint32_t r[3][4];
  for (int q = 0; q < 3; q++) {
    r[q][0] = slot;
    r[q][1] = step;
    r[q][2] = stream;
    r[q][3] = (uint32_t)q;
    philox4x32_10(r[q], k0, k1);
  }
  InflowIon o;
  const uint32_t* ur[6] = {&r[0][3], &r[1][0], &r[1][1], &r[1][2], &r[1][3], &r[2][0]};
  for (int a = 0; a < 3; a++) {
    double u1 = 1.0 - u24(*ur[2 * a]);  // (0,1]
    double u2 = u24(*ur[2 * a + 1]);
    o.v[a] = (a == 0 ? d.vsw : 0.0) + d.vth * sqrt(-log(u1)) * cos(PI_REF * u2);
  }
  double t0[3];
  t0[0] = (1.0 - u24(r[0][0])) * (double)d.L - (double)d.L;  // (-L, 0], exact
  const uint32_t col = slot / d.spc;  // slots of one y-z column are consecutive
  t0[1] = (double)(col % (uint32_t)d.nc[1]) + (1.0 - u24(r[0][1]));
  t0[2] = (double)(col / (uint32_t)d.nc[1]) + (1.0 - u24(r[0][2]));
  for (int a = 0; a < 3; a++) {
    o.th[a] = fma(o.v[a], d.sh[a], t0[a]);
    o.te[a] = fma(o.v[a], d.sh[a], o.th[a]);
  }
  for (int a = 1; a < 3; a++) {  // periodic y and z
    o.th[a] = wrap_open(o.th[a], (double)d.nc[a]);
    o.te[a] = wrap_open(o.te[a], (double)d.nc[a]);
  }
  const double ncx = (double)d.nc[0];
  o.dep = o.th[0] > 0.0 && o.th[0] <= ncx;
  o.keep = o.te[0] > 0.0 && o.te[0] <= ncx && o.th[0] <= ncx;
  return o;
}
// right-closed split of a cell coordinate t in (0, nc]
__device__ __forceinline__ int split_t(double t, double& f) {
  double i = ceil(t) - 1.0;
  f = t - i;
  return (int)i;
}

__global__ void k_inflow_count(uint32_t nslot, InflowDev d, uint32_t step, uint32_t k0,
                               uint32_t k1, uint32_t stream, uint32_t* cnt) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nslot) return;
  cnt[q] = inflow_slot(d, q, step, k0, k1, stream).keep ? 1u : 0u;
}

// One workgroup takes the slots of a patch of IN_PY x IN_PZ face columns.
// The half-step deposits of its entrants land within a few cells of the
// face and of the patch, so they are summed in an LDS window (x nodes
// [0, XW), y and z nodes from the patch start minus the thermal reach) and
// the window is added to the moments once; a corner outside the window
// (or every corner, when the window exceeds IN_WIN values) goes straight to
// HBM.  Fixed-point sums make the result independent of that split.
constexpr int IN_PY = 4, IN_PZ = 4, IN_WIN = 4096;
struct InflowWin {
  int XW, WY, WZ, ey, ez;
  bool on;
};
__device__ __forceinline__ void in_win_add(unsigned long long* a, int i, float v, int bits) {
  atomicAdd(a + i, to_fix(v, bits));
}
__device__ __forceinline__ void in_win_add(double* a, int i, double v, int) { atomicAdd(a + i, v); }

template <typename R>
__global__ __launch_bounds__(256) void k_inflow_fill(
    Geo g, InflowWin win, int64_t n0, PDev<R> P, InflowDev d, uint32_t step, uint32_t k0,
    uint32_t k1, uint32_t stream, const uint32_t* off, float bpv, float* bp, float* mr,
    uint32_t* id, uint32_t id0, typename AccOf<R>::T* acc) {
  using A = typename AccOf<R>::T;
  __shared__ A lacc[IN_WIN];
  const int npy = (d.nc[1] + IN_PY - 1) / IN_PY;
  const int py0 = (int)(blockIdx.x % npy) * IN_PY, pz0 = (int)(blockIdx.x / npy) * IN_PZ;
  const int by = py0 - win.ey, bz = pz0 - win.ez;  // window origin in y and z
  const int nwin = win.on ? win.XW * win.WY * win.WZ * 4 : 0;
  const int nwv = nwin / 4;
  for (int i = threadIdx.x; i < nwin; i += blockDim.x) lacc[i] = (A)0;
  __syncthreads();
  const uint32_t nps = (uint32_t)(IN_PY * IN_PZ) * d.spc;
  for (uint32_t sidx = threadIdx.x; sidx < nps; sidx += blockDim.x) {
    const uint32_t cp = sidx / d.spc;
    const int jy = py0 + (int)(cp % IN_PY), jz = pz0 + (int)(cp / IN_PY);
    if (jy >= d.nc[1] || jz >= d.nc[2]) continue;
    const uint32_t q = ((uint32_t)jz * (uint32_t)d.nc[1] + (uint32_t)jy) * d.spc + (sidx - cp * d.spc);
    const InflowIon o = inflow_slot(d, q, step, k0, k1, stream);
    if (o.dep) {  // the pass's moments at x^{n+1/2}: 1/beta_p and v/beta_p
      double f[3];
      int ijk[3];
      for (int a = 0; a < 3; a++) ijk[a] = split_t(o.th[a], f[a]);
      R w[8];
      pm::weights8((R)f[0], (R)f[1], (R)f[2], w);
      const R rr = (R)1 / (R)bpv;
      const R qm[4] = {rr, rr * (R)o.v[0], rr * (R)o.v[1], rr * (R)o.v[2]};
      const int dy = pm::wrapc(ijk[1] - by, g.ncy), dz = pm::wrapc(ijk[2] - bz, g.ncz);
      const bool inw = win.on && ijk[0] + 1 < win.XW && dy + 1 < win.WY && dz + 1 < win.WZ;
#pragma unroll
      for (int cc = 0; cc < 8; cc++) {
        const int cx = cc & 1, cy = (cc >> 2) & 1, cz = (cc >> 1) & 1;
        if (inw) {
          const int wi = ((dz + cz) * win.WY + dy + cy) * win.XW + ijk[0] + cx;
#pragma unroll
          for (int m = 0; m < 4; m++) in_win_add(lacc, wi + m * nwv, w[cc] * qm[m], fix_bits(m));
        } else {
          int64_t node = gidx(g, node_of_x(g, ijk[0] + cx), node_of(ijk[1] + cy, g.ncy),
                              node_of(ijk[2] + cz, g.ncz));
#pragma unroll
          for (int m = 0; m < 4; m++)
            acc_add(acc, node + (int64_t)m * g.nn, w[cc] * qm[m], fix_bits(m));
        }
      }
    }
    if (o.keep) {
      double f[3];
      int ijk[3];
      for (int a = 0; a < 3; a++) ijk[a] = split_t(o.te[a], f[a]);
      int64_t p = n0 + off[q];
      P.cell[p] = pack_cell(ijk[0], ijk[1], ijk[2]);
      P.fx[p] = (R)f[0]; P.fy[p] = (R)f[1]; P.fz[p] = (R)f[2];
      P.vx[p] = (R)o.v[0]; P.vy[p] = (R)o.v[1]; P.vz[p] = (R)o.v[2];
      P.v1x[p] = (R)o.v[0]; P.v1y[p] = (R)o.v[1]; P.v1z[p] = (R)o.v[2];
      if (bp) bp[p] = bpv;
      if (mr) mr[p] = 1.0f;
      if (id) id[p] = id0 + off[q];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nwv; i += blockDim.x) {
    const int ux = i % win.XW, uy = (i / win.XW) % win.WY, uz = i / (win.XW * win.WY);
    const int64_t node = gidx(g, node_of_x(g, ux), node_of(by + uy, g.ncy), node_of(bz + uz, g.ncz));
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const A v = lacc[i + m * nwv];
      if (v != (A)0) atomicAdd(acc + node + (int64_t)m * g.nn, v);
    }
  }
}

// The inflow ions' velocities lie in [lo, hi] per axis (the Box-Muller draw
// has u1 >= 2^-24, so |v - vsw xhat| <= 4.0794 vth).  Folded into the bounds
// P3 reduced (the grid's, and those of the tiles in the first ntxi tile
// columns along x, which hold every cell an entering ion reaches), they keep
// the next P3's windows valid.
struct VBox {
  float lo[3], hi[3];
};
__global__ void k_inflow_bounds(int* vbn, int ntx, int ntxi, int ntyz, VBox b) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > ntxi * ntyz) return;
  int* t = q == ntxi * ntyz ? vbn  // the grid, or tile (q % ntxi, q / ntxi)
                            : vbn + 6 * ((int64_t)ntx * (q / ntxi) + q % ntxi + 1);
  for (int a = 0; a < 3; a++) {
    atomicMin(t + a, pm::fkey(b.lo[a]));
    atomicMax(t + 3 + a, pm::fkey(b.hi[a]));
  }
}

// ---- FP64 reductions ----
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// particle sums: [0] charge, [1..3] momentum, [4] kinetic energy
template <typename R>
__global__ void k_diag_particles(int64_t n, PDev<R> P, double beta, double mion, double* out) {
  double s[5] = {0, 0, 0, 0, 0};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    double bp = P.bp ? (double)P.bp[p] : 1.0;
    double mr = P.mr ? (double)P.mr[p] : 1.0;
    double w = 1.0 / (beta * bp);
    double m = mion / mr;
    double vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
    s[0] += w;
    s[1] += w * m * vx;
    s[2] += w * m * vy;
    s[3] += w * m * vz;
    s[4] += w * 0.5 * m * (vx * vx + vy * vy + vz * vz) * 1e6;  // (km/s)^2 -> J
  }
  for (int q = 0; q < 5; q++) {
    double t = wave_sum(s[q]);
    if (__lane_id() == 0) atomicAdd(out + q, t);
  }
}

// mass-free kinetic energy scale * sum 1/2 |v|^2 / (mrat beta_p) (km/s)^2 ->
// (m/s)^2 of particles [0, n) of P into *out; Energy_diag's Evp
// (part_init.f:66-71) without the ion mass, with the weight 1/(beta beta_p)
// (scale = 1/beta, or -1/beta for ions that leave)
template <typename R>
__global__ void k_ke_sum(int64_t n, PDev<R> P, double scale, double* out) {
  double s = 0;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    double bp = P.bp ? (double)P.bp[p] : 1.0;
    double mr = P.mr ? (double)P.mr[p] : 1.0;
    double vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
    s += 0.5e6 * (vx * vx + vy * vy + vz * vz) / (mr * bp);
  }
  s = wave_sum(s);
  if (__lane_id() == 0) atomicAdd(out, s * scale);
}

// Energy_diag's grid sums (part_init.f:55-72) over the unique nodes:
// [0..2] sum b1_c^2, [3] sum |b1 - b0|^2, [4] sum |E|^2, [5] sum np |up|^2
template <typename R>
__global__ void k_energy_grid(Geo g, const double* b, const R* E, const R* up, const R* np,
                              double b0x, double b0y, double b0z, double* out) {
  double s[6] = {0, 0, 0, 0, 0, 0};
  const double b0[3] = {b0x, b0y, b0z};
  int64_t ncell = (int64_t)g.ncx * g.ncy * g.ncz;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ncell;
       q += (int64_t)gridDim.x * blockDim.x) {
    int i = 1 + (int)(q % g.ncx), j = 1 + (int)((q / g.ncx) % g.ncy),
        k = 1 + (int)(q / ((int64_t)g.ncx * g.ncy));
    int64_t o = gidx(g, i, j, k);
    double n = (double)np[o];
    for (int m = 0; m < 3; m++) {
      double bc = b[o + m * g.nn], e = (double)E[o + m * g.nn], u = (double)up[o + m * g.nn];
      s[m] += bc * bc;
      s[3] += (bc - b0[m]) * (bc - b0[m]);
      s[4] += e * e;
      s[5] += n * u * u;
    }
  }
  for (int m = 0; m < 6; m++) {
    double t = wave_sum(s[m]);
    if (__lane_id() == 0) atomicAdd(out + m, t);
  }
}

// grid sums over the unique dual cells: [5] sum (div b1)^2, [6] sum |b1|^2
__global__ void k_diag_grid(Geo g, const double* b, double dx, double dy, double dz,
                            double* out) {
  double s0 = 0, s1 = 0;
  int64_t ncell = (int64_t)g.ncx * g.ncy * g.ncz;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ncell;
       q += (int64_t)gridDim.x * blockDim.x) {
    int i = 1 + (int)(q % g.ncx), j = 1 + (int)((q / g.ncx) % g.ncy),
        k = 1 + (int)(q / ((int64_t)g.ncx * g.ncy));
    int64_t o = gidx(g, i, j, k);
    double d = (b[gidx(g, i + 1, j, k)] - b[o]) / dx +
               (b[gidx(g, i, j + 1, k) + g.nn] - b[o + g.nn]) / dy +
               (b[gidx(g, i, j, k + 1) + 2 * g.nn] - b[o + 2 * g.nn]) / dz;
    s0 += d * d;
    s1 += b[o] * b[o] + b[o + g.nn] * b[o + g.nn] + b[o + 2 * g.nn] * b[o + 2 * g.nn];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (__lane_id() == 0) {
    atomicAdd(out + 5, s0);
    atomicAdd(out + 6, s1);
  }
}

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

}  // namespace

#define SDISPATCH(c, ...)   \
  do {                      \
    if ((c)->prec == 64) {  \
      using R = double;     \
      __VA_ARGS__;          \
    } else {                \
      using R = float;      \
      __VA_ARGS__;          \
    }                       \
  } while (0)

int pl_load_abs(swap_ctx* c, int64_t n, const double* xp_d, const double* v_d,
                const double* v1_d) {
  if (n == 0) return 0;
  SDISPATCH(c, hipLaunchKernelGGL((k_load_abs<R>), dim3(nblk(n, 256)), dim3(256), 0, c->stream, n,
                                  (int64_t)0, pdev<R>(c), xp_d, v_d, v1_d, c->cfg.dx, c->cfg.dy,
                                  c->cfg.dz, c->cfg.ncx, c->cfg.ncy, c->cfg.ncz));
  return post(c);
}

int pl_set_abs(swap_ctx* c, const double* xp_d) {
  if (c->n == 0) return 0;
  SDISPATCH(c, hipLaunchKernelGGL((k_set_abs<R>), dim3(nblk(c->n, 256)), dim3(256), 0, c->stream,
                                  c->n, pdev<R>(c), xp_d, c->cfg.dx, c->cfg.dy, c->cfg.dz,
                                  c->cfg.ncx, c->cfg.ncy, c->cfg.ncz));
  return post(c);
}

int pl_export_abs(swap_ctx* c, double* xp_d) {
  if (c->n == 0) return 0;
  SDISPATCH(c, hipLaunchKernelGGL((k_export_abs<R>), dim3(nblk(c->n, 256)), dim3(256), 0,
                                  c->stream, c->n, pdev<R>(c), xp_d, c->cfg.dx, c->cfg.dy,
                                  c->cfg.dz));
  return post(c);
}

int pl_maxwellian_full(swap_ctx* c, int64_t off, int64_t n, double vsw, double vth, float bpv,
                       float mrv, uint32_t stream) {
  if (n == 0) return 0;
  uint32_t k0 = (uint32_t)c->cfg.seed, k1 = (uint32_t)(c->cfg.seed >> 32);
  SDISPATCH(c, hipLaunchKernelGGL((k_maxwellian<R>), dim3(nblk(n, 256)), dim3(256), 0, c->stream,
                                  n, off, pdev<R>(c), k0, k1, stream, vsw, vth, c->cfg.ncx,
                                  c->cfg.ncy, c->cfg.ncz, bpv, mrv, c->bp, c->mr, c->id));
  return post(c);
}

int sr_ionize_sw(swap_ctx* c, int step, double ndot, double vsw, double beta_pu,
                 int64_t* injected) {
  // dNi = ndot dt beta beta_pu dx dy delz nz (chem_rates.f:551), with the
  // stochastic round-up of :553-558 and nint of :560
  int nz_ref = c->cfg.ncz + 2;
  double dNi = ndot * c->cfg.dt * c->cfg.beta * beta_pu * c->cfg.dx * c->cfg.dy * c->cfg.dz *
               (double)nz_ref;
  if (dNi < 1.0) {
    uint32_t r[4] = {(uint32_t)step, 0u, 0x10u, 0u};
    philox4x32_10(r, (uint32_t)c->cfg.seed, (uint32_t)(c->cfg.seed >> 32));
    if (dNi > u24(r[0])) dNi = 1.0;
  }
  int64_t m = (int64_t)llround(dNi);
  if (m < 0) m = 0;
  if (c->n + m > c->cap && c->dead_pending) {  // room held by ions that left: drop them
    if (int e = so_sort(c)) return e;
  }
  if (c->n + m > c->cap) return set_err(c, SWAP_ECAP, "ionize: capacity %lld exceeded",
                                        (long long)c->cap);
  if (m > 0 && !(c->cfg.flags & SWAP_SPECIES) && beta_pu != 1.0)
    return set_err(c, SWAP_EINVAL, "ionize: beta_pu != 1 needs SWAP_SPECIES");
  uint32_t k0 = (uint32_t)c->cfg.seed, k1 = (uint32_t)(c->cfg.seed >> 32);
  if (m > 0) {
    SDISPATCH(c, hipLaunchKernelGGL((k_ionize_sw<R>), dim3(nblk(m, 256)), dim3(256), 0, c->stream,
                                    m, c->n, pdev<R>(c), k0, k1, 0x20000u + (uint32_t)step, vsw,
                                    c->cfg.ncx, c->cfg.ncy, c->cfg.ncz, (float)beta_pu, c->bp,
                                    c->mr, c->id, (uint32_t)c->next_id));
    if (int e = post(c)) return e;
    if (int e = sr_input_add(c, c->n, m, 1.0)) return e;
  }
  c->n += m;
  c->next_id += m;
  c->injected += m;
  *injected = m;
  return 0;
}

int sr_diagnostics(swap_ctx* c, swap_diag* d) {
  HIPCHK(c, hipMemsetAsync(c->dred, 0, 8 * sizeof(double), c->stream));
  if (c->n > 0)
    SDISPATCH(c, hipLaunchKernelGGL((k_diag_particles<R>), dim3(1024), dim3(256), 0, c->stream,
                                    c->n, pdev<R>(c), c->cfg.beta, 1.67e-27, c->dred));
  hipLaunchKernelGGL(k_diag_grid, dim3(1024), dim3(256), 0, c->stream, c->g, c->b1, c->cfg.dx,
                     c->cfg.dy, c->cfg.dz, c->dred);
  if (int e = post(c)) return e;
  double h[8];
  unsigned long long cnt[6];
  HIPCHK(c, hipMemcpyAsync(h, c->dred, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(cnt, c->dcnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d->charge = h[0];
  d->momentum[0] = h[1]; d->momentum[1] = h[2]; d->momentum[2] = h[3];
  d->kinetic = h[4];
  d->divb_l2 = sqrt(h[5]);
  d->b_energy = h[6];
  d->nparticles = c->n;
  d->escaped = c->escaped;
  d->injected = c->injected;
  d->window_misses = (int64_t)cnt[0];
  d->shrunk_windows = (int64_t)cnt[5];
  d->sorts = c->sorts;
  return 0;
}

namespace {

// ---- spherical neutral cloud: photoionization and charge exchange ----
// Neutral density n_n(r) = N0 (r0/r)^2 exp(-r/H), with r held at >= rmin
// (the 1/r^2 form of Ionize_pluto_mp, chem_rates.f:351, with the exp(-r/H)
// photoionization shell of chem_rates.f:114-116).
struct CloudDev {
  double cx, cy, cz, N0, r0, H, rmin, sigma, beta_min;
  double rate;  // vol beta dt / tau_photo: new ions per unit neutral density
  float mrat;
  int max_per_cell;
};
__device__ __forceinline__ double cloud_nn(const CloudDev& cl, double x, double y, double z) {
#pragma clang fp contract(off)  // the same roundings as the CPU restatement
  double r = sqrt((x - cl.cx) * (x - cl.cx) + (y - cl.cy) * (y - cl.cy) +
                  (z - cl.cz) * (z - cl.cz));
  r = r > cl.rmin ? r : cl.rmin;
  double a = cl.r0 / r;
  return cl.N0 * a * a * exp(-r / cl.H);
}

// per interior node (i fastest): the node's new-ion count and weight.  The
// expected number npofr = vol beta n_n dt / tau (chem_rates.f:352); at least
// one: nint(npofr) ions (:354); below one: one ion with probability npofr
// (:391).  More than max_per_cell become max_per_cell heavier ions
// (beta_p = max/count), so the deposited charge is unchanged.
__global__ void k_cloud_count(Geo g, double dx, double dy, double dz, CloudDev cl, uint32_t k0,
                              uint32_t k1, uint32_t stream, uint32_t* cnt, float* bp) {
#pragma clang fp contract(off)
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t nint_nodes = (int64_t)g.ncx * g.ncy * g.ncz;
  if (q >= nint_nodes) return;
  int i = 1 + (int)(q % g.ncx), j = 1 + (int)((q / g.ncx) % g.ncy),
      k = 1 + (int)(q / ((int64_t)g.ncx * g.ncy));
  double npofr = cl.rate * cloud_nn(cl, (i + 1) * dx, (j + 1) * dy, (k + 1) * dz);
  uint32_t n;
  if (npofr >= 1.0) {
    n = npofr < 4.0e9 ? (uint32_t)llround(npofr) : 4000000000u;
  } else {
    uint32_t r[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32), stream, 0xFFFFFFFFu};
    philox4x32_10(r, k0, k1);
    n = npofr > u24(r[0]) ? 1u : 0u;
  }
  float w = 1.0f;
  if (n > (uint32_t)cl.max_per_cell) {
    // m heavier ions carry the n: beta_p = m/n, with m raised so beta_p >= beta_min
    double m = (double)cl.max_per_cell, lo = ceil(cl.beta_min * (double)n);
    if (lo > m) m = lo < (double)n ? lo : (double)n;
    w = (float)(m / (double)n);
    n = (uint32_t)m;
  }
  cnt[q] = n;
  bp[q] = w;
}

// the new ions of each node at rest, uniform in the node's dual cell
// (xp = q(i) + (u - 0.5) dx, chem_rates.f:363-365), mrat = 1/m_pu
template <typename R>
__global__ void k_cloud_fill(Geo g, int64_t n0, PDev<R> P, CloudDev cl, uint32_t k0, uint32_t k1,
                             uint32_t stream, const uint32_t* cnt, const uint32_t* off,
                             const float* bpn, float* bp, float* mr, uint32_t* id,
                             uint32_t id0) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t nint_nodes = (int64_t)g.ncx * g.ncy * g.ncz;
  if (q >= nint_nodes) return;
  const int s[3] = {1 + (int)(q % g.ncx), 1 + (int)((q / g.ncx) % g.ncy),
                    1 + (int)(q / ((int64_t)g.ncx * g.ncy))};
  const int nc[3] = {g.ncx, g.ncy, g.ncz};
  for (uint32_t l = 0; l < cnt[q]; l++) {
    uint32_t r[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32), stream, l};
    philox4x32_10(r, k0, k1);
    int ijk[3];
    double f[3];
    for (int a = 0; a < 3; a++) {
      double t = (double)s[a] + u24(r[a]) - 0.5;  // cell coordinate, in (0.5, nc + 0.5)
      // past the last face: the periodic image, or on an open x the mirror
      // image inside the domain
      if (t > (double)nc[a]) t = a == 0 && g.ox ? 2.0 * nc[a] - t : t - (double)nc[a];
      double ci = ceil(t) - 1.0;
      ijk[a] = (int)ci;
      f[a] = t - ci;
    }
    int64_t p = n0 + off[q] + l;
    P.cell[p] = pack_cell(ijk[0], ijk[1], ijk[2]);
    P.fx[p] = (R)f[0]; P.fy[p] = (R)f[1]; P.fz[p] = (R)f[2];
    P.vx[p] = (R)0; P.vy[p] = (R)0; P.vz[p] = (R)0;
    P.v1x[p] = (R)0; P.v1y[p] = (R)0; P.v1z[p] = (R)0;
    bp[p] = bpn[q];
    mr[p] = cl.mrat;
    if (id) id[p] = id0 + off[q] + l;
  }
}

// res_chex (chem_rates.f:805-843) with the cloud's density at the ion:
// probability dt n_n sigma |v|; a hit stops the ion (v = 0) and makes it a
// cloud ion (mrat = 1/m_pu)
template <typename R>
__global__ void k_chex(int64_t n, PDev<R> P, const uint32_t* id, double dx, double dy,
                       double dz, double dt, CloudDev cl, uint32_t k0, uint32_t k1,
                       uint32_t stream, float* mr, unsigned long long* count) {
#pragma clang fp contract(off)
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t c = P.cell[p];
  if (c == DEAD_CELL) return;  // left through an open x face, dropped by the next sort
  double x = ((double)cell_i(c) + 1.0 + (double)P.fx[p]) * dx;
  double y = ((double)cell_j(c) + 1.0 + (double)P.fy[p]) * dy;
  double z = ((double)cell_k(c) + 1.0 + (double)P.fz[p]) * dz;
  double vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
  double prob = dt * cloud_nn(cl, x, y, z) * cl.sigma * sqrt(vx * vx + vy * vy + vz * vz);
  // the draw belongs to the ion: keyed by its id when ids are kept, so it
  // does not depend on where the sort has put the ion
  const uint64_t who = id ? (uint64_t)id[p] : (uint64_t)p;
  uint32_t r[4] = {(uint32_t)who, (uint32_t)(who >> 32), stream, 0u};
  philox4x32_10(r, k0, k1);
  if (u24(r[0]) < prob) {
    P.vx[p] = (R)0; P.vy[p] = (R)0; P.vz[p] = (R)0;
    mr[p] = cl.mrat;
    atomicAdd(count, 1ull);
  }
}

}  // namespace

int sr_ionize_cloud(swap_ctx* c, int step, const swap_cloud* cl, int64_t* injected,
                    int64_t* chex) {
  *injected = 0;
  *chex = 0;
  if (!(c->cfg.flags & SWAP_SPECIES))
    return set_err(c, SWAP_EINVAL, "ionize_cloud needs SWAP_SPECIES (beta_p, mrat arrays)");
  if (!(cl->N0 >= 0) || !(cl->r0 > 0) || !(cl->H > 0) || !(cl->tau_photo > 0) ||
      !(cl->rmin > 0) || cl->max_per_cell < 1 || !(cl->mrat_pu > 0) || !(cl->sigma_chex >= 0) ||
      !(cl->beta_min >= 0 && cl->beta_min <= 1))
    return set_err(c, SWAP_EINVAL, "ionize_cloud: bad cloud parameters");
  const Geo& g = c->g;
  const int64_t nin = (int64_t)g.ncx * g.ncy * g.ncz;
  if (!c->ccnt) {
    HIPCHK(c, hipMalloc(&c->ccnt, 4 * (size_t)(nin + 1)));
    HIPCHK(c, hipMalloc(&c->coff, 4 * (size_t)(nin + 1)));
    HIPCHK(c, hipMalloc(&c->cbp, 4 * (size_t)(nin + 1)));
  }
  CloudDev d;
  d.cx = cl->cx; d.cy = cl->cy; d.cz = cl->cz;
  d.N0 = cl->N0; d.r0 = cl->r0; d.H = cl->H; d.rmin = cl->rmin; d.sigma = cl->sigma_chex;
  d.beta_min = cl->beta_min;
  d.rate = c->cfg.dx * c->cfg.dy * c->cfg.dz * c->cfg.beta * c->cfg.dt / cl->tau_photo;
  d.mrat = (float)cl->mrat_pu;
  d.max_per_cell = cl->max_per_cell;
  const uint32_t k0 = (uint32_t)c->cfg.seed, k1 = (uint32_t)(c->cfg.seed >> 32);
  // charge exchange of the ions present (before this step's new ions)
  if (c->n > 0 && cl->sigma_chex > 0) {
    HIPCHK(c, hipMemsetAsync(c->dcnt + 3, 0, sizeof(unsigned long long), c->stream));
    SDISPATCH(c, hipLaunchKernelGGL((k_chex<R>), dim3(nblk(c->n, 256)), dim3(256), 0, c->stream,
                                    c->n, pdev<R>(c), (const uint32_t*)c->id, c->cfg.dx,
                                    c->cfg.dy, c->cfg.dz,
                                    c->cfg.dt, d, k0, k1, 0x40000u + (uint32_t)step, c->mr,
                                    c->dcnt + 3));
    if (int e = post(c)) return e;
    unsigned long long h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, c->dcnt + 3, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *chex = (int64_t)h;
  }
  // photoionization: counts per node, offsets, then the new ions
  hipLaunchKernelGGL(k_cloud_count, dim3(nblk(nin, 256)), dim3(256), 0, c->stream, g, c->cfg.dx,
                     c->cfg.dy, c->cfg.dz, d, k0, k1, 0x38000u + (uint32_t)step, c->ccnt,
                     c->cbp);
  if (int e = post(c)) return e;
  if (int e = so_scan_u32(c, nin, c->ccnt, c->coff)) return e;
  uint32_t tot = 0;
  HIPCHK(c, hipMemcpyAsync(&tot, c->coff + nin, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->n + (int64_t)tot > c->cap && c->dead_pending) {  // room held by ions that left
    if (int e = so_sort(c)) return e;
  }
  if (c->n + (int64_t)tot > c->cap)
    return set_err(c, SWAP_ECAP, "ionize_cloud: %lld + %u ions exceed capacity %lld",
                   (long long)c->n, tot, (long long)c->cap);
  if (tot > 0) {
    SDISPATCH(c, hipLaunchKernelGGL((k_cloud_fill<R>), dim3(nblk(nin, 256)), dim3(256), 0,
                                    c->stream, g, c->n, pdev<R>(c), d, k0, k1,
                                    0x30000u + (uint32_t)step, c->ccnt, c->coff, c->cbp, c->bp,
                                    c->mr, c->id, (uint32_t)c->next_id));
    if (int e = post(c)) return e;
  }
  c->n += tot;
  c->next_id += tot;
  c->injected += tot;
  *injected = tot;
  return 0;
}

namespace {
// order-preserving int image of a float, for integer atomic min/max
__device__ __forceinline__ int fkey_s(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ int wave_mini(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_maxi(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
// signed per-axis velocity bounds: out[0..2] min keys, out[3..5] max keys
template <typename R>
__global__ void k_vbounds(int64_t n, PDev<R> P, int* out) {
  int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int hi[3] = {-0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    int k[3] = {fkey_s((float)P.vx[p]), fkey_s((float)P.vy[p]), fkey_s((float)P.vz[p])};
    for (int a = 0; a < 3; a++) {
      lo[a] = min(lo[a], k[a]);
      hi[a] = max(hi[a], k[a]);
    }
  }
  for (int a = 0; a < 3; a++) {
    int l = wave_mini(lo[a]), h = wave_maxi(hi[a]);
    if (__lane_id() == 0) {
      atomicMin(out + a, l);
      atomicMax(out + 3 + a, h);
    }
  }
}
// empties the grid's bounds and every tile's (a pass then sizes its windows
// from the grid's bounds until a P3 has reduced the tiles')
__global__ void k_vbounds_init(int64_t nvb, int* out) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < nvb) out[q] = (q % 6) < 3 ? 0x7fffffff : -0x7fffffff - 1;
}
}  // namespace

int sr_vbounds(swap_ctx* c) {
  hipLaunchKernelGGL(k_vbounds_init, dim3(nblk(vb_ints(c), 256)), dim3(256), 0, c->stream,
                     vb_ints(c), c->vb);
  if (c->n > 0)
    SDISPATCH(c, hipLaunchKernelGGL((k_vbounds<R>), dim3(1024), dim3(256), 0, c->stream, c->n,
                                    pdev<R>(c), c->vb));
  return post(c);
}

int sr_set_inflow(swap_ctx* c, double ppc, double vsw, double vth, double beta_p) {
  if (!c->g.ox) return set_err(c, SWAP_EINVAL, "set_inflow needs SWAP_OUTFLOW_X");
  if (!(ppc >= 0) || !(vth >= 0) || !std::isfinite(vsw) || !(beta_p > 0))
    return set_err(c, SWAP_EINVAL, "set_inflow: bad parameters");
  if (beta_p != 1.0 && !(c->cfg.flags & SWAP_SPECIES))
    return set_err(c, SWAP_EINVAL, "set_inflow: beta_p != 1 needs SWAP_SPECIES");
  // the sampler's speeds along x lie within vsw -+ vth sqrt(ln 2^24) = 4.08 vth
  double reach = (fabs(vsw) + 4.08 * vth) * c->cfg.dt / c->cfg.dx;
  if (!(reach < 1048576.0)) return set_err(c, SWAP_EINVAL, "set_inflow: slab too deep");
  int L = std::max(1, (int)ceil(reach));
  // round(ppc L) slots per y-z column of the face, the columns' slots consecutive
  double ns = std::round(ppc * (double)L) * (double)c->cfg.ncy * (double)c->cfg.ncz;
  if (!(ns < 4.0e9)) return set_err(c, SWAP_EINVAL, "set_inflow: too many slots");
  c->in_ppc = ppc;
  c->in_vsw = vsw;
  c->in_vth = vth;
  c->in_bp = beta_p;
  c->in_L = L;
  c->in_slots = (int64_t)ns;
  return 0;
}

int sr_inflow_step(swap_ctx* c) {
  c->in_last = 0;
  const uint32_t nslot = (uint32_t)c->in_slots;
  if (!c->g.ox || nslot == 0) return 0;
  InflowDev d;
  d.vsw = c->in_vsw;
  d.vth = c->in_vth;
  d.sh[0] = 0.5 * c->cfg.dt / c->cfg.dx;
  d.sh[1] = 0.5 * c->cfg.dt / c->cfg.dy;
  d.sh[2] = 0.5 * c->cfg.dt / c->cfg.dz;
  d.L = c->in_L;
  d.nc[0] = c->cfg.ncx; d.nc[1] = c->cfg.ncy; d.nc[2] = c->cfg.ncz;
  d.spc = (uint32_t)(nslot / ((uint64_t)c->cfg.ncy * (uint64_t)c->cfg.ncz));
  if ((int64_t)nslot > c->islots) {
    for (uint32_t** b : {&c->icnt, &c->ioff, &c->iblk}) {
      if (*b) HIPCHK(c, hipFree(*b));
      *b = nullptr;
    }
    c->islots = 0;
    HIPCHK(c, hipMalloc(&c->icnt, 4 * ((size_t)nslot + 1)));
    HIPCHK(c, hipMalloc(&c->ioff, 4 * ((size_t)nslot + 1)));
    HIPCHK(c, hipMalloc(&c->iblk, 4 * ((size_t)nslot / 4096 + 4)));
    c->islots = nslot;
  }
  const uint32_t k0 = (uint32_t)c->cfg.seed, k1 = (uint32_t)(c->cfg.seed >> 32);
  const uint32_t stream = 0x50000u, step = (uint32_t)c->step_count;
  hipLaunchKernelGGL(k_inflow_count, dim3(nblk(nslot, 256)), dim3(256), 0, c->stream, nslot, d,
                     step, k0, k1, stream, c->icnt);
  if (int e = post(c)) return e;
  if (int e = so_scan_u32(c, nslot, c->icnt, c->ioff, c->iblk)) return e;
  uint32_t tot = 0;
  HIPCHK(c, hipMemcpyAsync(&tot, c->ioff + nslot, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->n + (int64_t)tot > c->cap)
    return set_err(c, SWAP_ECAP, "inflow: %lld + %u ions exceed capacity %lld",
                   (long long)c->n, tot, (long long)c->cap);
  // deposit window of a patch: x nodes up to the half-step reach (th <= L/2
  // cells), y and z nodes of the patch plus the thermal reach of a half move
  InflowWin win;
  win.XW = (c->in_L + 1) / 2 + 2;
  win.ey = (int)std::ceil(4.08 * c->in_vth * d.sh[1]);
  win.ez = (int)std::ceil(4.08 * c->in_vth * d.sh[2]);
  win.WY = IN_PY + 2 * win.ey + 1;
  win.WZ = IN_PZ + 2 * win.ez + 1;
  win.on = (int64_t)win.XW * win.WY * win.WZ * 4 <= IN_WIN;
  const unsigned npatch = (unsigned)(((c->cfg.ncy + IN_PY - 1) / IN_PY) *
                                     ((c->cfg.ncz + IN_PZ - 1) / IN_PZ));
  SDISPATCH(c, hipLaunchKernelGGL((k_inflow_fill<R>), dim3(npatch), dim3(256), 0, c->stream,
                                  c->g, win, c->n, pdev<R>(c), d, step, k0, k1, stream, c->ioff,
                                  (float)c->in_bp, c->bp, c->mr, c->id, (uint32_t)c->next_id,
                                  (typename AccOf<R>::T*)c->macc));
  if (int e = post(c)) return e;
  if (int e = sr_input_add(c, c->n, tot, 1.0)) return e;
  c->n += tot;
  c->next_id += tot;
  c->injected += tot;
  c->in_last = tot;
  const double a = 4.08 * c->in_vth;
  const double lo[3] = {c->in_vsw - a, -a, -a}, hi[3] = {c->in_vsw + a, a, a};
  VBox b;
  for (int q = 0; q < 3; q++) {  // one float step outward covers the rounding to R
    b.lo[q] = std::nextafter((float)lo[q], -INFINITY);
    b.hi[q] = std::nextafter((float)hi[q], INFINITY);
  }
  // entering ions end the step in cells [0, L) of x
  const int ntyz = c->ntile[1] * c->ntile[2];
  const int ntxi = std::min(c->ntile[0], (c->in_L + TILE_X - 1) / TILE_X);
  hipLaunchKernelGGL(k_inflow_bounds, dim3(nblk(ntxi * ntyz + 1, 256)), dim3(256), 0, c->stream,
                     c->vb_next, c->ntile[0], ntxi, ntyz, b);
  return post(c);
}

// the input energy account (Energy_diag's input_E, global.f:47): ions that
// enter (loads, sources, inflow) add their kinetic energy, ions that leave
// through an open x subtract theirs; mass-free, on the device
int sr_input_add(swap_ctx* c, int64_t off, int64_t n, double sign) {
  if (n <= 0) return 0;
  double* acc = c->dred + DRED_INPUT;
  SDISPATCH(c, {
    PDev<R> P = pdev<R>(c);
    P.cell += off;
    for (R** a : {&P.fx, &P.fy, &P.fz, &P.vx, &P.vy, &P.vz, &P.v1x, &P.v1y, &P.v1z}) *a += off;
    if (P.bp) P.bp += off;
    if (P.mr) P.mr += off;
    hipLaunchKernelGGL((k_ke_sum<R>), dim3((unsigned)std::min<int64_t>(1024, nblk(n, 256))),
                       dim3(256), 0, c->stream, n, P, sign / c->cfg.beta, acc);
  });
  return post(c);
}
int sr_input_reset(swap_ctx* c) {
  HIPCHK(c, hipMemsetAsync(c->dred + DRED_INPUT, 0, sizeof(double), c->stream));
  return 0;
}

// Energy_diag (part_init.f:14-139) in SI units with the reference's constants
// (inputs.f:19,50-57): over the unique nodes (the reference sums the i = 1
// ghost layer too and one y plane, j = 2, of its quasi-1-D grid) and with
// the weight 1/(beta beta_p) for every ion (the reference divides by beta
// and multiplies by beta_p, part_init.f:69-70)
int sr_energy_diag(swap_ctx* c, double mion, const double b0[3], swap_energy* d) {
  double* r = c->dred + DRED_ENERGY;  // [0] sum KE/m, [1..6] grid sums
  HIPCHK(c, hipMemsetAsync(r, 0, 8 * sizeof(double), c->stream));
  if (c->n > 0)
    SDISPATCH(c, hipLaunchKernelGGL((k_ke_sum<R>), dim3(1024), dim3(256), 0, c->stream, c->n,
                                    pdev<R>(c), 1.0 / c->cfg.beta, r));
  SDISPATCH(c, hipLaunchKernelGGL((k_energy_grid<R>), dim3(1024), dim3(256), 0, c->stream, c->g,
                                  c->b1, (const R*)c->fld[SWAP_F_E], (const R*)c->fld[SWAP_F_UP],
                                  (const R*)c->fld[SWAP_F_NP], b0[0], b0[1], b0[2], r + 1));
  if (int e = post(c)) return e;
  double h[8], in = 0;
  HIPCHK(c, hipMemcpyAsync(h, r, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&in, c->dred + DRED_INPUT, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double q = 1.6e-19, pi = 3.14159, mu0 = pi * 4.0e-7, eps0 = 8.85e-12, km = 1e3;
  const double m_q = mion / q, vol = c->cfg.dx * c->cfg.dy * c->cfg.dz * km * km * km;
  const double bfac = vol / (2.0 * mu0) * m_q * m_q;
  d->Evp = mion * h[0];
  d->EB1x = bfac * h[1];
  d->EB1y = bfac * h[2];
  d->EB1z = bfac * h[3];
  d->EB1 = bfac * h[4];
  d->EE = eps0 * vol / 2.0 * (m_q * km) * (m_q * km) * h[5];
  d->Euf = 0.5 * mion * vol / (km * km * km) * km * km * h[6];
  d->input_E = mion * in;
  d->total_E = d->Evp + d->EE + d->EB1;
  d->norm_E = d->input_E != 0 ? d->total_E / d->input_E : 0.0;
  return 0;
}