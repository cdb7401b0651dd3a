// swap_common.h -- shared device/host definitions of the gfx950 library.
//
// Grid storage is the reference's Fortran layout with two ghost layers:
// index (i,j,k) 0-based storage, i fastest, component-major for vectors.
// Storage index 0 and n-1 along an axis are the ghosts that boundary.f's
// periodic() fills with copies of n-2 and 1 (boundary.f:18-46).  Every device
// kernel writes ghosts as well as interior nodes by evaluating the interior
// formula at the "canonical" node, which is exactly the value periodic()
// would copy there, so device arrays are always periodic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SWAP_HD __host__ __device__ __forceinline__

struct Geo {
  int nx, ny, nz;     // storage extents including ghosts
  int ncx, ncy, ncz;  // physical cells = nx-2 ...
  int64_t nn;         // nx*ny*nz
  int ox;             // x inflow/outflow instead of periodic
};

SWAP_HD int64_t gidx(const Geo& g, int i, int j, int k) {
  return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * k);
}
// canonical storage index of a (possibly ghost) storage index s in [0,n)
SWAP_HD int canon(int s, int n) { return s == 0 ? n - 2 : (s == n - 1 ? 1 : s); }
// storage index of unwrapped node coordinate u (node u = lower corner of
// cell u); folds ghost contributions into their periodic images
// (update_np's fold, gutsp.f:1496-1498)
// r mod nc in [0, nc) for any r.  Callers pass coordinates within a period
// of the grid, which take one conditional add and one conditional subtract;
// anything farther (never in the particle passes) takes the division in a
// branch of its own, so the common path carries no division (a division
// costs some 20 instructions per call)
SWAP_HD int wrap_range(int r, int nc) {
  r += r < 0 ? nc : 0;
  r -= r >= nc ? nc : 0;
  if (__builtin_expect(r < 0 || r >= nc, 0)) {
    r %= nc;
    if (r < 0) r += nc;
  }
  return r;
}
SWAP_HD int node_of(int u, int nc) { return wrap_range(u - 1, nc) + 1; }

// The x axis of an inflow/outflow run (Geo::ox) has no periodic image: its
// ghost layers are zero-gradient copies of the nearest interior node, and
// deposits on the inflow face node (storage 0) fold onto the first interior
// node.  canon_x/node_of_x are canon/node_of for x under that rule.
SWAP_HD int canon_x(const Geo& g, int s) {
  return g.ox ? (s < 1 ? 1 : (s > g.nx - 2 ? g.nx - 2 : s)) : canon(s, g.nx);
}
SWAP_HD int node_of_x(const Geo& g, int u) {
  return g.ox ? (u < 1 ? 1 : (u > g.ncx ? g.ncx : u)) : node_of(u, g.ncx);
}
// Density of an open x node: the particle domain is cells 0..ncx-1; node 1
// collects the whole of cell 0 (through the fold above) and half of cell 1,
// node ncx half of cell ncx-1.  Dividing their deposits by those shares (1.5
// and 0.5 cells) keeps a uniform plasma uniform up to both faces.
SWAP_HD double face_scale(const Geo& g, int s) {
  return !g.ox ? 1.0 : (s == 1 ? 2.0 / 3.0 : (s == g.ncx ? 2.0 : 1.0));
}
// cell of a particle that left through an open x face; the sort that
// follows removes it
constexpr uint32_t DEAD_CELL = 0xFFFFFFFFu;

// packed cell: i | j<<11 | k<<21 (0-based cell indices)
SWAP_HD uint32_t pack_cell(int i, int j, int k) {
  return (uint32_t)i | ((uint32_t)j << 11) | ((uint32_t)k << 21);
}
SWAP_HD int cell_i(uint32_t c) { return (int)(c & 2047u); }
SWAP_HD int cell_j(uint32_t c) { return (int)((c >> 11) & 1023u); }
SWAP_HD int cell_k(uint32_t c) { return (int)(c >> 21); }

// Philox-4x32-10 (Salmon et al., SC'11), counter-based: output depends only
// on (counter, key), so every particle draws its own stream.
SWAP_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// uniform in [0,1) with 24 random bits (exact in float and double)
SWAP_HD double u24(uint32_t r) { return (double)(r >> 8) * (1.0 / 16777216.0); }

// ---- moment accumulation ----
// FP32 deposits are summed exactly: each FP32 contribution (a particle's
// product w q, or an FP32 partial sum over a few neighbouring particles) is
// rounded to a 64-bit integer count of 2^-bits units and added with integer
// atomics, in LDS and in HBM.  Integer addition is exact and associative, so
// the node sums carry no accumulation error and do not depend on the order
// in which atomics are served; the only rounding is that of each
// contribution to the fixed grid: 2^-24 for the density moment (weights
// times 1/beta_p), 2^-15 km/s for the velocity moments -- below 1e-9 of the
// node sums of the workloads here, and below the FP32 rounding of a partial
// sum over a few particles.  A contribution below 2^31 units (|w r| < 128,
// |w r v| < 65536 km/s) converts in a few operations; larger ones take an
// exact 64-bit path.  The units are as coarse as that allows so that the
// partial sums over the runs of a sorted wavefront (up to 64 particles at
// several hundred km/s) stay on the short path.  The FP64 build accumulates
// doubles.
constexpr int FIX_BITS_N = 24, FIX_BITS_V = 15;
SWAP_HD constexpr int fix_bits(int moment) { return moment == 0 ? FIX_BITS_N : FIX_BITS_V; }
template <typename R>
struct AccOf {
  using T = unsigned long long;  // FP32 build: fixed point
};
template <>
struct AccOf<double> {
  using T = double;
};
__device__ __noinline__ unsigned long long to_fix_wide(float x) {
  return (unsigned long long)(long long)x;
}
// round(v 2^bits) as a two's-complement 64-bit integer
__device__ __forceinline__ unsigned long long to_fix(float v, int bits) {
  const float x = rintf(v * __int_as_float((127 + bits) << 23));
  if (__builtin_expect(fabsf(x) < 2147483648.0f, 1))
    return (unsigned long long)(long long)(int)x;
  return to_fix_wide(x);
}
__device__ __forceinline__ double from_fix(unsigned long long a, int bits) {
  return (double)(long long)a * (double)__int_as_float((127 - bits) << 23);
}
__device__ __forceinline__ void acc_add(unsigned long long* a, int64_t i, float v, int bits) {
  if (v != 0.f) atomicAdd(a + i, to_fix(v, bits));
}
__device__ __forceinline__ void acc_add(double* a, int64_t i, double v, int) {
  if (v != 0.0) atomicAdd(a + i, v);
}
__device__ __forceinline__ double acc_read(const unsigned long long* a, int64_t i, int bits) {
  return from_fix(a[i], bits);
}
__device__ __forceinline__ double acc_read(const double* a, int64_t i, int) { return a[i]; }
