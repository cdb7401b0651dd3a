// k_pass.hip -- the three fused particle passes of swap_step.
//
//   P1: moments at x^n of r and r(1.5 v - 0.5 v1)             (maind.f:385-399)
//   P2: Ep, v+-, deposit r (v+ + v-)/2 at x^n                   (maind.f:400-402)
//   P3: Ep, v+-, v1 <- v, v <- v+ + dt/2 Ep, half move to x^{n+1/2}, deposit
//       r and r v there, second half move to x^{n+1}            (maind.f:404-413,479)
// with r = 1/beta_p.  The division of update_up by np(node)
// (gutsp.f:1704-1756) is done once per node after the sums.
//
// One workgroup runs one work item: a run of one tile's particles after the
// counting sort, processed in chunks of PASS_WG particles, one per lane.
// Each lane forms its particle's 8 corners x M moments; runs of lanes whose
// particles share a cell (the sort leaves each tile's particles in a few
// runs in cell order) are summed across the wavefront by a segmented scan,
// so that only a run's last lane adds into the LDS window.  Window and HBM
// additions are exact fixed-point integer atomics (FP64 in the FP64 build).
// The window covers the item's deposit footprint and is added to HBM once
// per item.  The window of P1/P2 is the item's cell box at x^n; P3's is that
// box widened by the half-step displacement range from signed per-axis
// velocity bounds of the item's tile (reduced by the previous P3).
// Particles outside the window deposit straight to HBM (counted as misses).
//
// The gather of the 6-component node source S = (ajc - upc, btc) in P2 and
// P3 reads one 32-byte record per corner through the caches.
#include <type_traits>
#include "particle_math.h"

using namespace pm;

namespace {

template <int MODE>
struct NMom {
  static constexpr int value = MODE == 2 ? 3 : 4;
};

// nearest periodic image of cell c to a box centre
__device__ __forceinline__ int unwrap(int c, int centre, int nc) {
  int d = c - centre;
  if (2 * d > nc) c -= nc;
  else if (2 * d < -nc) c += nc;
  return c;
}

// ---- moment accumulation into the LDS window ----
// FP32 build: a pass multiplies each particle's moments by 2^bits of their
// fixed-point unit once (exact: a power of two), so its FP32 products and
// run sums are the unit counts themselves, rounded exactly as the products
// and sums of the unscaled moments would be.  Adding a sum is then rounding
// to an integer and converting: two instructions for a sum below 2^30
// units; in general the 32-bit form or, for |sum| >= 2^31, an exact split
// into two 32-bit halves, selected per value without a branch.  The FP64
// build adds doubles.
constexpr float TWO31F = 2147483648.f, TWO32F = 4294967296.f;
// |x| < 2^30: floor(x + 1/2) in one instruction (v_cvt_rpi_i32_f32, whose
// add is exact: it equals floor(x + 1/2) taken in FP64 on every input
// checked, ties rounding up where rintf goes to even), then the sign
// extension.  Any rounding that depends on x alone keeps the sums exact and
// order-free; this one spends two instructions per value instead of three.
__device__ __forceinline__ unsigned long long fix_small(float x) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return (unsigned long long)(long long)r;
}
__device__ __forceinline__ unsigned long long fix_any(float x) {  // |rint(x)| < 2^62
  const float r = rintf(x);
  const unsigned long long small = (unsigned long long)(long long)(int)r;
  // |r| >= 2^31: r is a multiple of ulp(r) >= 2^8, and so is the low half
  // r - hi 2^32 in [0, 2^32): it has at most 24 significant bits and the fma
  // is exact.  (Below 2^31 the low half of a negative r may need more bits
  // than FP32 holds; the 32-bit form covers those.)
  const float hi = floorf(r * (1.f / TWO32F));
  const float lo = fmaf(hi, -TWO32F, r);
  const unsigned long long wide =
      ((unsigned long long)(unsigned)(int)hi << 32) | (unsigned long long)(unsigned)lo;
  return fabsf(r) < TWO31F ? small : wide;
}
// one contribution straight to the HBM accumulators (a window miss)
__device__ __forceinline__ void gacc_add(unsigned long long* a, int64_t i, float v) {
  if (v != 0.f) atomicAdd(a + i, fix_any(v));
}
__device__ __forceinline__ void gacc_add(double* a, int64_t i, double v) {
  if (v != 0.0) atomicAdd(a + i, v);
}
// A lane's 8 corners x M sums into the window at node `key`.  The corners
// (the reference's order, gutsp.f:1422-1429) are key + {0, 1} + {0, sz} +
// {0, sy} (sy, sz: the window's y and z strides) and moment m lies WS values
// further on (a compile-time stride), so every add is one DS atomic at one
// of four base addresses plus an immediate offset.  mag bounds |sum| of
// every value (FP32 build); below 2^30 the two-instruction conversion
// holds for all 8 x M of them, with one branch.
template <int M, int WS>
__device__ __forceinline__ void win_put(unsigned long long* lacc, int key, int sy, int sz,
                                        const float (&a)[8][M], float mag) {
  unsigned long long* b[4] = {lacc + key, lacc + key + sz, lacc + key + sy, lacc + key + sy + sz};
  if (mag < 1073741824.f) {
#pragma unroll
    for (int cc = 0; cc < 8; cc++)
#pragma unroll
      for (int m = 0; m < M; m++) atomicAdd(b[cc >> 1] + (cc & 1) + m * WS, fix_small(a[cc][m]));
  } else {
#pragma unroll
    for (int cc = 0; cc < 8; cc++)
#pragma unroll
      for (int m = 0; m < M; m++) atomicAdd(b[cc >> 1] + (cc & 1) + m * WS, fix_any(a[cc][m]));
  }
}
template <int M, int WS>
__device__ __forceinline__ void win_put(double* lacc, int key, int sy, int sz,
                                        const double (&a)[8][M], float) {
  double* b[4] = {lacc + key, lacc + key + sz, lacc + key + sy, lacc + key + sy + sz};
#pragma unroll
  for (int cc = 0; cc < 8; cc++)
#pragma unroll
    for (int m = 0; m < M; m++) atomicAdd(b[cc >> 1] + (cc & 1) + m * WS, a[cc][m]);
}

// ---- segmented inclusive scan over the 64 lanes of a wavefront ----
// Lane L belongs to the run starting at lane s(L).  Hillis-Steele inside each
// row of 16 lanes (row_shr 1,2,4,8), then the row totals cross rows with
// row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3); a contribution is
// taken only from a lane of the same run.  A step whose condition holds in
// no lane adds nothing and is skipped (lv, the same in every lane): after a
// drift the runs are a few lanes long and most waves need only the first
// shifts.
struct SegCond {
  bool c1, c2, c4, c8, c15, c31;
  unsigned lv;  // bit b: step b is needed by some lane
};
__device__ __forceinline__ SegCond seg_cond(int L, int s) {
  SegCond c;
  c.c1 = L - 1 >= s;
  c.c2 = L - 2 >= s;
  c.c4 = L - 4 >= s;
  c.c8 = L - 8 >= s;
  c.c15 = ((L >> 4) & 1) && s <= (L & ~15) - 1;
  c.c31 = (L >> 5) && s <= 31;
  c.lv = __builtin_amdgcn_readfirstlane(
      (__ballot(c.c1) ? 1u : 0u) | (__ballot(c.c2) ? 2u : 0u) | (__ballot(c.c4) ? 4u : 0u) |
      (__ballot(c.c8) ? 8u : 0u) | (__ballot(c.c15) ? 16u : 0u) | (__ballot(c.c31) ? 32u : 0u));
  return c;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dppf(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xf, true));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dppf(double v) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xf, true);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
// One step of the scan on every value of a lane's sums (and on mag).  FP32:
// a <- t * f + a with t the neighbour's value read through DPP and f 1
// where the step's condition holds in this lane, else 0: one v_fmac_f32_dpp
// per value (t * 1 + a rounds once, exactly as a + t; t * 0 + a is a).
// The compiler keeps the DPP read a separate move, so the step is written
// out.  The hazard recognizer does not look into the block: its first
// instruction may read a register that the instruction before it wrote, or
// follow a VALU write of EXEC, so it opens with 5 wait states; inside the
// block no value is read by DPP within 2 instructions of its write.  Rows
// outside the row mask are not written (the neighbour lies outside them).
#define SWAP_FMAC_DPP(i, CTL) "v_fmac_f32_dpp %" #i ", %" #i ", %[f] " CTL "\n"
#define SWAP_FMAC_DPP8(r, i0, i1, i2, i3, i4, i5, i6, i7, CTL)                        \
  SWAP_FMAC_DPP(i0, CTL) SWAP_FMAC_DPP(i1, CTL) SWAP_FMAC_DPP(i2, CTL)                 \
  SWAP_FMAC_DPP(i3, CTL) SWAP_FMAC_DPP(i4, CTL) SWAP_FMAC_DPP(i5, CTL)                 \
  SWAP_FMAC_DPP(i6, CTL) SWAP_FMAC_DPP(i7, CTL)
#define SWAP_SEG_STEP_ASM(NAME, CTL)                                                          \
  __device__ __forceinline__ void NAME(float (&a)[8][4], float& mag, float f) {              \
    asm("s_nop 4\n" SWAP_FMAC_DPP8(0, 0, 1, 2, 3, 4, 5, 6, 7, CTL)                          \
            SWAP_FMAC_DPP8(1, 8, 9, 10, 11, 12, 13, 14, 15, CTL)                             \
                SWAP_FMAC_DPP8(2, 16, 17, 18, 19, 20, 21, 22, 23, CTL)                       \
                    SWAP_FMAC_DPP8(3, 24, 25, 26, 27, 28, 29, 30, 31, CTL)                   \
                        SWAP_FMAC_DPP(32, CTL)                                               \
        : "+&v"(a[0][0]), "+&v"(a[0][1]), "+&v"(a[0][2]), "+&v"(a[0][3]), "+&v"(a[1][0]),    \
          "+&v"(a[1][1]), "+&v"(a[1][2]), "+&v"(a[1][3]), "+&v"(a[2][0]), "+&v"(a[2][1]),    \
          "+&v"(a[2][2]), "+&v"(a[2][3]), "+&v"(a[3][0]), "+&v"(a[3][1]), "+&v"(a[3][2]),    \
          "+&v"(a[3][3]), "+&v"(a[4][0]), "+&v"(a[4][1]), "+&v"(a[4][2]), "+&v"(a[4][3]),    \
          "+&v"(a[5][0]), "+&v"(a[5][1]), "+&v"(a[5][2]), "+&v"(a[5][3]), "+&v"(a[6][0]),    \
          "+&v"(a[6][1]), "+&v"(a[6][2]), "+&v"(a[6][3]), "+&v"(a[7][0]), "+&v"(a[7][1]),    \
          "+&v"(a[7][2]), "+&v"(a[7][3]), "+&v"(mag)                                         \
        : [f] "v"(f));                                                                       \
  }                                                                                          \
  __device__ __forceinline__ void NAME(float (&a)[8][3], float& mag, float f) {              \
    asm("s_nop 4\n" SWAP_FMAC_DPP8(0, 0, 1, 2, 3, 4, 5, 6, 7, CTL)                          \
            SWAP_FMAC_DPP8(1, 8, 9, 10, 11, 12, 13, 14, 15, CTL)                             \
                SWAP_FMAC_DPP8(2, 16, 17, 18, 19, 20, 21, 22, 23, CTL)                       \
                    SWAP_FMAC_DPP(24, CTL)                                                   \
        : "+&v"(a[0][0]), "+&v"(a[0][1]), "+&v"(a[0][2]), "+&v"(a[1][0]), "+&v"(a[1][1]),    \
          "+&v"(a[1][2]), "+&v"(a[2][0]), "+&v"(a[2][1]), "+&v"(a[2][2]), "+&v"(a[3][0]),    \
          "+&v"(a[3][1]), "+&v"(a[3][2]), "+&v"(a[4][0]), "+&v"(a[4][1]), "+&v"(a[4][2]),    \
          "+&v"(a[5][0]), "+&v"(a[5][1]), "+&v"(a[5][2]), "+&v"(a[6][0]), "+&v"(a[6][1]),    \
          "+&v"(a[6][2]), "+&v"(a[7][0]), "+&v"(a[7][1]), "+&v"(a[7][2]), "+&v"(mag)         \
        : [f] "v"(f));                                                                       \
  }
SWAP_SEG_STEP_ASM(seg_shr1, "row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1")
SWAP_SEG_STEP_ASM(seg_shr2, "row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1")
SWAP_SEG_STEP_ASM(seg_shr4, "row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1")
SWAP_SEG_STEP_ASM(seg_shr8, "row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1")
SWAP_SEG_STEP_ASM(seg_bc15, "row_bcast:15 row_mask:0xa bank_mask:0xf bound_ctrl:1")
SWAP_SEG_STEP_ASM(seg_bc31, "row_bcast:31 row_mask:0xc bank_mask:0xf bound_ctrl:1")
#undef SWAP_SEG_STEP_ASM
#undef SWAP_FMAC_DPP8
#undef SWAP_FMAC_DPP
// FP64: the neighbour's value where the condition holds
template <int CTRL, int ROWS, int M>
__device__ __forceinline__ void seg_step(double (&a)[8][M], float& mag, bool cond) {
#pragma unroll
  for (int cc = 0; cc < 8; cc++)
#pragma unroll
    for (int m = 0; m < M; m++) {
      double t = dppf<CTRL, ROWS>(a[cc][m]);
      a[cc][m] = cond ? a[cc][m] + t : a[cc][m];
    }
  const float t = dppf<CTRL, ROWS>(mag);
  mag = cond ? mag + t : mag;
}
// the run sums a and their magnitude bound mag
template <int M>
__device__ __forceinline__ void segscan(float (&a)[8][M], float& mag, const SegCond& c) {
  if (c.lv & 1u) seg_shr1(a, mag, c.c1 ? 1.f : 0.f);
  if (c.lv & 2u) seg_shr2(a, mag, c.c2 ? 1.f : 0.f);
  if (c.lv & 4u) seg_shr4(a, mag, c.c4 ? 1.f : 0.f);
  if (c.lv & 8u) seg_shr8(a, mag, c.c8 ? 1.f : 0.f);
  if (c.lv & 16u) seg_bc15(a, mag, c.c15 ? 1.f : 0.f);
  if (c.lv & 32u) seg_bc31(a, mag, c.c31 ? 1.f : 0.f);
}
template <int M>
__device__ __forceinline__ void segscan(double (&a)[8][M], float& mag, const SegCond& c) {
  if (c.lv & 1u) seg_step<0x111, 0xf>(a, mag, c.c1);
  if (c.lv & 2u) seg_step<0x112, 0xf>(a, mag, c.c2);
  if (c.lv & 4u) seg_step<0x114, 0xf>(a, mag, c.c4);
  if (c.lv & 8u) seg_step<0x118, 0xf>(a, mag, c.c8);
  if (c.lv & 16u) seg_step<0x142, 0xa>(a, mag, c.c15);
  if (c.lv & 32u) seg_step<0x143, 0xc>(a, mag, c.c31);
}

// The sums a of each lane's cell `cur` (-1: nothing to add) into the window:
// runs of consecutive lanes with the same cell are summed across the
// wavefront first (sorted particles), and only each run's last lane adds
// into the window; lanes that share no cell with a neighbour add directly.
// FH: lanes that start a run whatever their neighbour holds (bit L: lane
// L), so that no run crosses them and the scan steps that reach across
// them are skipped, at the price of one more tail (a few more active lanes
// in the same window atomics) per crossing.  Every lane of the wavefront
// calls it.
template <int M, int WS, unsigned long long FH, typename T, typename A>
__device__ __forceinline__ void put_runs(T (&a)[8][M], float& mag, int cur, A* lacc, int sy,
                                         int sz, int lane) {
  const int prevkey = __shfl_up(cur, 1);
  const unsigned long long hm = __ballot(lane == 0 || prevkey != cur) | FH;
  if (hm != ~0ull) {
    const unsigned long long le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
    const SegCond sc = seg_cond(lane, 63 - __clzll(hm & le));
    segscan(a, mag, sc);
  }
  const bool tail = lane == 63 || ((hm >> (lane + 1)) & 1ull);
  if (tail && cur >= 0) win_put<M, WS>(lacc, cur, sy, sz, a, mag);
}

// The value of lane (lane ^ J) for every lane of the wavefront, in VALU:
// DPP quad permutes (J = 1, 2), row shifts into alternate 4- or 8-lane banks
// (J = 4, 8), and the row-pair and half swaps of gfx950 (J = 16, 32).
template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t x, int lane) {
  if constexpr (J == 1) {
    return __builtin_amdgcn_update_dpp(0u, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (J == 2) {
    return __builtin_amdgcn_update_dpp(0u, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (J == 4) {
    const uint32_t o = __builtin_amdgcn_update_dpp(0u, x, 0x104, 0xF, 0x5, false);  // row_shl:4
    return __builtin_amdgcn_update_dpp(o, x, 0x114, 0xF, 0xA, false);               // row_shr:4
  } else if constexpr (J == 8) {
    const uint32_t o = __builtin_amdgcn_update_dpp(0u, x, 0x108, 0xF, 0x3, false);  // row_shl:8
    return __builtin_amdgcn_update_dpp(o, x, 0x118, 0xF, 0xC, false);               // row_shr:8
  } else if constexpr (J == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return ((lane >> 4) & 1) ? r[0] : r[1];
  } else {
    static_assert(J == 32, "xor distance");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return lane >= 32 ? r[0] : r[1];
  }
}
// one compare-exchange of a bitonic network: blocks of K lanes ascending
// where (lane & K) == 0, the lower lane of each pair J apart keeping the min
template <int J, int K>
__device__ __forceinline__ uint32_t bitonic_ce(uint32_t v, int lane) {
  const uint32_t o = xor_lane<J>(v, lane);
  const bool keep_min = ((lane & K) == 0) == ((lane & J) == 0);
  return keep_min ? min(v, o) : max(v, o);
}
// ascending sort of 64 distinct values, one per lane (21 compare-exchanges)
__device__ __forceinline__ uint32_t wave_sort(uint32_t v, int lane) {
  v = bitonic_ce<1, 2>(v, lane);
  v = bitonic_ce<2, 4>(v, lane);  v = bitonic_ce<1, 4>(v, lane);
  v = bitonic_ce<4, 8>(v, lane);  v = bitonic_ce<2, 8>(v, lane);  v = bitonic_ce<1, 8>(v, lane);
  v = bitonic_ce<8, 16>(v, lane); v = bitonic_ce<4, 16>(v, lane);
  v = bitonic_ce<2, 16>(v, lane); v = bitonic_ce<1, 16>(v, lane);
  v = bitonic_ce<16, 32>(v, lane); v = bitonic_ce<8, 32>(v, lane);
  v = bitonic_ce<4, 32>(v, lane);  v = bitonic_ce<2, 32>(v, lane); v = bitonic_ce<1, 32>(v, lane);
  v = bitonic_ce<32, 64>(v, lane); v = bitonic_ce<16, 64>(v, lane);
  v = bitonic_ce<8, 64>(v, lane);  v = bitonic_ce<4, 64>(v, lane);
  v = bitonic_ce<2, 64>(v, lane);  v = bitonic_ce<1, 64>(v, lane);
  return v;
}

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}


// node of unwrapped x node coordinate u: periodic, or on an open x (OX) the
// clamp of node_of_x.  A template flag, so that the periodic kernels carry
// none of the open-x code (their registers set their occupancy).
template <bool OX>
__device__ __forceinline__ int xnode(const Geo& g, int u) {
  return OX ? (u < 1 ? 1 : (u > g.ncx ? g.ncx : u)) : node_of(u, g.ncx);
}

// One particle per lane in every pass.  Occupancy moves in whole 512-thread
// workgroups (2 waves on each of a CU's 4 SIMDs): two of them (4 waves per
// SIMD) share a CU when a wave needs at most 128 VGPRs and the LDS holds
// both windows (2 x 81.6 KB of the CU's 160 KB).  The FP32 passes are held
// to that (P3 fits 127 VGPRs with no spills); P3 at one workgroup per CU
// (165 VGPRs when its divergent exits kept the 32 sums live on every path)
// took 9.8 ms against 7.7.  The FP64 build runs one workgroup per CU with
// the registers it needs (2 is the compiler's own default for 512 threads).
// Measured losers: 4 consecutive particles per lane summed in registers
// (P1/P2 on the step right after a sort; one workgroup per CU) once the
// sort leaves cells in short runs, 5.1 / 5.2 ms against 4.8 / 4.3 ms for one
// particle per lane; P3 depositing the next step's P1 moments at x^{n+1}
// into a second window (1024-thread workgroups) instead of P1 reading the
// particles again: P3 7.45 -> 12.5 ms against P1's 5.2 ms.
template <typename R>
struct PassWaves {
  static constexpr int value = sizeof(R) == 4 ? 4 : 2;
};

template <typename R, int MODE, bool OX>
__global__ __launch_bounds__(PASS_WG, (PassWaves<R>::value)) void k_pass(
    Geo g, PDev<R> P, PCoef<R> K, const R* __restrict__ S,
    typename AccOf<R>::T* __restrict__ gacc, const Item* __restrict__ items,
    const int* __restrict__ d_nitems, int64_t n_sorted, int64_t n_total,
    const int* __restrict__ boxes, int* boxes_next, const int* __restrict__ vb, int* vb_next,
    unsigned long long* dcnt, const int* __restrict__ tboxes) {
  constexpr int M = NMom<MODE>::value;
  constexpr int GM0 = MODE == 2 ? 1 : 0;  // moment index of window moment 0
  constexpr int WS = WIN_VALUES / M;       // window values per moment (LDS stride)
  constexpr bool FIX = sizeof(R) == 4;     // fixed-point moment sums (FP32 build)
  using A = typename AccOf<R>::T;
  __shared__ A lacc[WIN_VALUES];
  __shared__ int sred[12];

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int nitems = *d_nitems;
  const int nc[3] = {g.ncx, g.ncy, g.ncz};
  int64_t i0, i1;
  int bl[3] = {0, 0, 0}, bh[3] = {-1, -1, -1};
  bool windowed = false;
  int tile = -1;
  const int w = blockIdx.x;
  const int* b;
  if (w < nitems) {
    Item it = items[w];
    i0 = it.start;
    i1 = it.end;
    tile = it.tile;
    b = boxes + 6 * w;
  } else {  // the unsorted tail (ions appended since the sort), with its own boxes
    int64_t t = w - nitems;
    i0 = n_sorted + t * TAIL_ITEM;
    if (i0 >= n_total) return;
    i1 = i0 + TAIL_ITEM < n_total ? i0 + TAIL_ITEM : n_total;
    b = tboxes + 6 * t;
  }
  {
    for (int a = 0; a < 3; a++) {
      bl[a] = b[a];
      bh[a] = b[3 + a];
      int cen = (bl[a] + bh[a]) >> 1;  // keep the box centre in [0, nc)
      int sh = (cen >= 0 ? cen / nc[a] : -((-cen + nc[a] - 1) / nc[a])) * nc[a];
      if (a == 0 && OX) sh = 0;  // an open x has no periodic images
      bl[a] -= sh;
      bh[a] -= sh;
    }
    windowed = bh[0] >= bl[0] && bh[1] >= bl[1] && bh[2] >= bl[2];
  }
  int lo_m[3] = {0, 0, 0}, hi_m[3] = {0, 0, 0};
  float mid[3] = {0.f, 0.f, 0.f};  // centre of each axis's displacement range
  if (MODE == 3 && windowed) {
    // the cell shift of a half move by d in [dlo, dhi] cells lies in
    // [floor(dlo), ceil(dhi)]; 10 % of the largest speed covers the change of
    // v within one step (rotation by Omega dt and the E x B acceleration).
    // The bounds are those of the item's tile at the previous P3 (the grid's
    // when the tile had no particles then), so that fast ions (pickup ions
    // near a source, say) widen only the windows of their own region; the
    // range is clamped to WIN_SHIFT cells.  Particles outside the window
    // deposit straight to HBM.
    const int* tb = vb + 6 * (tile + 1);
    if (!(tb[0] <= tb[3] && tb[1] <= tb[4] && tb[2] <= tb[5])) tb = vb;
    float vmax_abs = 0.f;
    for (int a = 0; a < 3; a++)
      vmax_abs = fmaxf(vmax_abs, fmaxf(-fval(tb[a]), fval(tb[3 + a])));
    float sl = 0.1f * vmax_abs;
    for (int a = 0; a < 3; a++) {
      float dlo = (float)K.s[a] * (fval(tb[a]) - sl), dhi = (float)K.s[a] * (fval(tb[3 + a]) + sl);
      if (!(dlo <= dhi)) windowed = false;  // unset bounds (NaN)
      else {
        const float cl = fminf(fmaxf(dlo, -WIN_SHIFT), WIN_SHIFT);
        const float ch = fminf(fmaxf(dhi, -WIN_SHIFT), WIN_SHIFT);
        lo_m[a] = (int)floorf(cl);
        hi_m[a] = (int)ceilf(ch);
        mid[a] = 0.5f * (cl + ch);
      }
    }
  }
  int u0[3], W[3], cen[3];
  for (int a = 0; a < 3; a++) {
    u0[a] = bl[a] + lo_m[a];
    W[a] = bh[a] + hi_m[a] - u0[a] + 2;
    cen[a] = (bl[a] + bh[a]) >> 1;
  }
  if (MODE == 3 && windowed) {
    // a P3 window over the LDS budget (boxes grown by a few steps of drift)
    // gives up cells of its shift range one at a time, the widest axis's
    // side farther from the range's centre first: the bulk keeps the window
    // and only the fastest particles deposit straight to HBM (counted in
    // dcnt[5] per item, beside the misses in dcnt[0])
    if (tid == 0 && (int64_t)W[0] * W[1] * W[2] * M > WIN_VALUES) atomicAdd(dcnt + 5, 1ull);
    for (int it = 0; it < 24 && (int64_t)W[0] * W[1] * W[2] * M > WIN_VALUES; it++) {
      int a = 0;
      for (int b = 1; b < 3; b++)
        if (hi_m[b] - lo_m[b] > hi_m[a] - lo_m[a]) a = b;
      if (hi_m[a] - lo_m[a] <= 0) break;
      if ((float)hi_m[a] - mid[a] >= mid[a] - (float)lo_m[a]) {
        hi_m[a]--;
      } else {
        lo_m[a]++;
        u0[a]++;
      }
      W[a]--;
    }
  }
  if (windowed) {
    // a box itself over the LDS budget (the thermal spread of several steps
    // since the sort) is clipped: its widest axis gives up one cell at a
    // time, at alternate ends, so that the dense core keeps the window and
    // the particles of the outer cells deposit straight to HBM (misses)
    int flip[3] = {0, 0, 0};
    for (int it = 0; it < 96 && (int64_t)W[0] * W[1] * W[2] * M > WIN_VALUES; it++) {
      int a = W[1] > W[0] ? 1 : 0;
      if (W[2] > W[a]) a = 2;
      if (W[a] <= 2) break;
      if (flip[a]) u0[a]++;
      W[a]--;
      flip[a] ^= 1;
    }
  }
  int64_t Wn64 = (int64_t)W[0] * W[1] * W[2];
  if (windowed && (W[0] < 2 || W[1] < 2 || W[2] < 2 || Wn64 * M > WIN_VALUES)) windowed = false;
  const int Wn = windowed ? (int)Wn64 : 0;
  for (int q = tid; q < Wn * M; q += PASS_WG) lacc[q % Wn + (q / Wn) * WS] = (A)0;
  if (MODE == 3 && tid < 12) sred[tid] = (tid % 6) < 3 ? 0x7fffffff : -0x7fffffff - 1;
  __syncthreads();

  const int sy = W[0], sz = W[0] * W[1];  // window strides of the y and z corners
  int bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int bmax[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
  float vlo[3] = {3e38f, 3e38f, 3e38f}, vhi[3] = {-3e38f, -3e38f, -3e38f};
  unsigned long long misses = 0;

  // Consecutive lanes take consecutive particles, so every load and store
  // instruction covers whole cache lines.  A chunk's particle reads are
  // issued before the previous chunk's work, so that their latency overlaps
  // it.  A lane past the item's end reads a loadable particle and, like a
  // lane holding an ion that left, contributes zero moments: the 8 x M sums
  // are plain products with no zeroing on divergent paths (each such path
  // cost 32 moves per chunk); only the stores and the out-of-window deposit
  // are guarded.
  struct Pre {
    uint32_t cell;
    R f[3], v[3], v1[3];
    float bp;
  } nxt{};
  auto fetch = [&](int64_t pb) {
    const int64_t p = pb < i1 ? pb : i0;
    nxt.cell = P.cell[p];
    nxt.f[0] = P.fx[p]; nxt.f[1] = P.fy[p]; nxt.f[2] = P.fz[p];
    nxt.v[0] = P.vx[p]; nxt.v[1] = P.vy[p]; nxt.v[2] = P.vz[p];
    if (MODE == 1) {
      nxt.v1[0] = P.v1x[p]; nxt.v1[1] = P.v1y[p]; nxt.v1[2] = P.v1z[p];
    }
    nxt.bp = P.bp ? P.bp[p] : 1.f;
  };
  if (i0 < i1) fetch(i0 + tid);
  for (int64_t chunk = i0; chunk < i1; chunk += PASS_WG) {
    const int64_t pb = chunk + tid;
    const int64_t p = pb < i1 ? pb : i0;
    const Pre cu = nxt;
    if (chunk + PASS_WG < i1) fetch(pb + PASS_WG);
    const bool valid = pb < i1 && cu.cell != DEAD_CELL;
    const uint32_t c = valid ? cu.cell : 0u;
    const int ci = cell_i(c), cj = cell_j(c), ck = cell_k(c);
    R f0 = cu.f[0], f1 = cu.f[1], f2 = cu.f[2];
    const R r = P.bp ? (R)1 / (R)cu.bp : (R)1;
    R w8[8], q[M], vn[3] = {0, 0, 0}, mr = (R)1;
    weights8(f0, f1, f2, w8);
    int ux = OX ? ci : unwrap(ci, cen[0], g.ncx), uy = unwrap(cj, cen[1], g.ncy),
        uz = unwrap(ck, cen[2], g.ncz);
    bool alive = valid;  // false once the particle left through an open x face
    const R v[3] = {cu.v[0], cu.v[1], cu.v[2]};
    if constexpr (MODE == 1) {
      q[0] = r;
      q[1] = r * ((R)1.5 * v[0] - (R)0.5 * cu.v1[0]);
      q[2] = r * ((R)1.5 * v[1] - (R)0.5 * cu.v1[1]);
      q[3] = r * ((R)1.5 * v[2] - (R)0.5 * cu.v1[2]);
    } else {
      R Ag[3], Bg[3], E[3], vm[3], vpl[3];
      gather6(g, S, gidx(g, ci, cj, ck), w8, Ag, Bg);
      if (P.mr) mr = (R)P.mr[p];
      ep_of(Ag, Bg, mr, E);
      rotate(K.dt, v, E, Bg, mr, vm, vpl);
      if constexpr (MODE == 2) {
#pragma unroll
        for (int m = 0; m < 3; m++) q[m] = r * ((R)0.5 * (vpl[m] + vm[m]));
      } else {
        const R hdt = (R)0.5 * K.dt;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          vn[m] = vpl[m] + hdt * E[m];
          const float fv = (float)vn[m];  // plain compares: no NaN quieting
          vlo[m] = valid && fv < vlo[m] ? fv : vlo[m];
          vhi[m] = valid && fv > vhi[m] ? fv : vhi[m];
        }
        ux += half_move(f0, vn[0], K.s[0]);
        uy += half_move(f1, vn[1], K.s[1]);
        uz += half_move(f2, vn[2], K.s[2]);
        if (OX) alive = alive && ux >= 0 && ux < g.ncx;  // gone before x^{n+1/2}
        weights8(f0, f1, f2, w8);
        q[0] = r; q[1] = r * vn[0]; q[2] = r * vn[1]; q[3] = r * vn[2];
      }
    }
    // fixed-point units (a power-of-two scale: exact), and the bound of
    // every |w q| of this particle
    float pmag = 0.f;
    if constexpr (FIX) {
#pragma unroll
      for (int m = 0; m < M; m++) {
        q[m] *= (R)__int_as_float((127 + fix_bits(GM0 + m)) << 23);
        pmag = fmaxf(pmag, fabsf((float)q[m]));
      }
    }
    const int lx = ux - u0[0], ly = uy - u0[1], lz = uz - u0[2];
    const bool inw = alive && windowed && lx >= 0 && ly >= 0 && lz >= 0 && lx < W[0] - 1 &&
                     ly < W[1] - 1 && lz < W[2] - 1;
    if (alive && !inw) {
      misses++;
#pragma unroll
      for (int cc = 0; cc < 8; cc++) {
        int64_t node = gidx(g, xnode<OX>(g, ux + (cc & 1)), node_of(uy + ((cc >> 2) & 1), g.ncy),
                            node_of(uz + ((cc >> 1) & 1), g.ncz));
#pragma unroll
        for (int m = 0; m < M; m++) gacc_add(gacc, node + (int64_t)m * g.nn, w8[cc] * q[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < M; m++) q[m] = inw ? q[m] : (R)0;
    const int cur = inw ? lx + W[0] * (ly + W[1] * lz) : -1;
    float amag = inw ? pmag : 0.f;
    R d0 = f0, d1 = f1, d2 = f2;  // fractions at the deposit position
    {
      // the deposit: this particle's 8 corners x M moments (in P3 before
      // the second half move and the stores, which it does not need:
      // 7.56 -> 7.50 ms with 128 VGPRs and a 16-byte spill, against the
      // deposit last)
      weights8(d0, d1, d2, w8);
      R acc[8][M];
#pragma unroll
      for (int cc = 0; cc < 8; cc++)
#pragma unroll
        for (int m = 0; m < M; m++) acc[cc][m] = w8[cc] * q[m];
      // runs cut at every 16 lanes in P1 (no row_bcast steps), every 8 in P2
      // (no row_shr:8 either), none in P3, whose deposit atomics already meet
      // more address conflicts: measured best of 8 / 16 / 32 / none per pass
      constexpr unsigned long long FH = MODE == 1   ? 0x0001000100010001ull
                                        : MODE == 2 ? 0x0101010101010101ull
                                                    : 0ull;
      put_runs<M, WS, FH>(acc, amag, cur, lacc, sy, sz, lane);
    }
    if constexpr (MODE == 3) {
      ux += half_move(f0, vn[0], K.s[0]);
      uy += half_move(f1, vn[1], K.s[1]);
      uz += half_move(f2, vn[2], K.s[2]);
      uint32_t cnew = 0u;
      if (valid) {
        if (OX && (!alive || ux < 0 || ux >= g.ncx)) {
          cnew = DEAD_CELL;  // the next sort drops it
        } else {
          cnew = pack_cell(wrapc(ux, g.ncx), wrapc(uy, g.ncy), wrapc(uz, g.ncz));
          bmin[0] = min(bmin[0], ux); bmax[0] = max(bmax[0], ux);
          bmin[1] = min(bmin[1], uy); bmax[1] = max(bmax[1], uy);
          bmin[2] = min(bmin[2], uz); bmax[2] = max(bmax[2], uz);
        }
      }
      // The wavefront's particles are stored in the order of their new
      // cells within the wavefront's own 64 slots, so that the next step's
      // P1 and P2 find each wavefront's particles in runs of equal cells
      // (the scan then merges every run into one window add).  Every array
      // of a particle moves with it.  The lanes inside the work item are a
      // prefix of the wavefront and sort before the others, so the slots
      // stay inside the item; on an open x the ions that left (before this
      // pass or during it) sort behind the live ones and keep the velocity,
      // weight and id that the escape account reads when a sort drops them.
      const bool inrange = pb < i1;
      const bool live = valid && cnew != DEAD_CELL;
      const uint32_t key = live ? (((uint32_t)uz & 255u) << 16) | (((uint32_t)uy & 255u) << 8) |
                                      ((uint32_t)ux & 255u)
                                : (inrange ? 0xFFFFFEu : 0xFFFFFFu);
      const uint32_t prev = (uint32_t)__shfl_up((int)key, 1);
      int rank = lane;
      if (__ballot(lane > 0 && prev > key) != 0ull) {
        // sort (key, lane) over the wavefront; lane L then holds the
        // element of sorted position L and tells that element its rank
        const uint32_t kv = wave_sort((key << 6) | (uint32_t)lane, lane);
        rank = __builtin_amdgcn_ds_permute((int)((kv & 63u) << 2), lane);
      }
      const int64_t dst = pb - lane + rank;
      if (inrange && dst < i1) {
        if (valid) {
          P.v1x[dst] = v[0]; P.v1y[dst] = v[1]; P.v1z[dst] = v[2];
          P.vx[dst] = vn[0]; P.vy[dst] = vn[1]; P.vz[dst] = vn[2];
          P.fx[dst] = f0; P.fy[dst] = f1; P.fz[dst] = f2;
          P.cell[dst] = cnew;
        } else if (dst != p) {  // an ion that left before this pass
          P.vx[dst] = v[0]; P.vy[dst] = v[1]; P.vz[dst] = v[2];
          P.cell[dst] = DEAD_CELL;
        }
        if (dst != p) {
          if (P.bp) {
            P.bp[dst] = (float)cu.bp;
            P.mr[dst] = (float)mr;
          }
          if (P.id) {
            const uint32_t pid = P.id[p];
            P.id[dst] = pid;
          }
        }
      }
    }
  }
  if (misses) atomicAdd(dcnt, misses);
  if (MODE == 3) {
    // the next step's windows: cell box of x^{n+1} and signed velocity bounds
    // (the grid's and the item's tile's)
    for (int a = 0; a < 3; a++) {
      int m0 = wave_min(bmin[a]), m1 = wave_max(bmax[a]);
      int k0 = wave_min(fkey(vlo[a])), k1 = wave_max(fkey(vhi[a]));
      if (lane == 0) {
        atomicMin(&sred[a], m0);
        atomicMax(&sred[3 + a], m1);
        atomicMin(&sred[6 + a], k0);
        atomicMax(&sred[9 + a], k1);
      }
    }
  }
  __syncthreads();
  if (MODE == 3 && tid < 6) {
    if (w < nitems) {
      if (tid == 0 && sred[0] <= sred[3] && sred[1] <= sred[4] && sred[2] <= sred[5] &&
          (int64_t)(sred[3] - sred[0] + 2) * (sred[4] - sred[1] + 2) * (sred[5] - sred[2] + 2) *
                  4 > BOX_SORT_VALUES)
        atomicAdd(dcnt + 4, 1ull);  // next step's P1 window would lose too much
      int* tn = vb_next + 6 * (tile + 1);
      if (tid < 3) {
        atomicMin(boxes_next + 6 * w + tid, sred[tid]);
        atomicMin(tn + tid, sred[6 + tid]);
      } else {
        atomicMax(boxes_next + 6 * w + tid, sred[tid]);
        atomicMax(tn + tid, sred[6 + tid]);
      }
    }
    if (tid < 3) atomicMin(vb_next + tid, sred[6 + tid]);
    else atomicMax(vb_next + tid, sred[6 + tid]);
  }
  for (int qn = tid; qn < Wn; qn += PASS_WG) {
    int lx = qn % W[0], ly = (qn / W[0]) % W[1], lz = qn / (W[0] * W[1]);
    int64_t node = gidx(g, xnode<OX>(g, u0[0] + lx), node_of(u0[1] + ly, g.ncy),
                        node_of(u0[2] + lz, g.ncz));
#pragma unroll
    for (int m = 0; m < M; m++) {
      A v = lacc[qn + m * WS];
      if (v != (A)0) atomicAdd(gacc + node + (int64_t)m * g.nn, v);
    }
  }
}

__global__ void k_pass_reset(int64_t nitems, int64_t nvb, int* boxes_next, int* vb_next,
                             unsigned long long* box_over) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q == 0) *box_over = 0;
  if (q < nvb) vb_next[q] = (q % 6) < 3 ? 0x7fffffff : -0x7fffffff - 1;
  if (q >= nitems * 6) return;
  boxes_next[q] = (q % 6) < 3 ? 0x7fffffff : -0x7fffffff;
}

// Cell boxes of the unsorted tail's work items (ions appended since the last
// sort: inflow entrants, cloud ions), TAIL_ITEM consecutive particles each:
// the entrants of one step fill the inflow face column by column (some 22
// columns of cells per item at 64 ions per cell and 2.9 cells of drift per
// step) and the cloud's ions node by node, so a tail item is compact and its
// deposits fit a window like a sorted item's, also after the thermal spread
// of the few steps until the next sort.  On a
// periodic axis the box is taken in whichever of two frames (the cells, or
// the cells shifted by half the axis) spans less, so an item across the
// periodic seam keeps a compact box (its low end may then be negative).
// Ions that left are skipped; an item with no live ion gets the empty box
// (no window).
__global__ __launch_bounds__(256) void k_tail_boxes(int64_t n_sorted, int64_t n_total,
                                                    const uint32_t* __restrict__ cell, Geo g,
                                                    int* tboxes) {
  __shared__ int red[12];  // lo, hi of the cells; lo, hi of the shifted cells
  if (threadIdx.x < 12) red[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0x7fffffff : -0x7fffffff;
  __syncthreads();
  const int nc[3] = {g.ncx, g.ncy, g.ncz};
  const int64_t i0 = n_sorted + (int64_t)blockIdx.x * TAIL_ITEM;
  const int64_t i1 = i0 + TAIL_ITEM < n_total ? i0 + TAIL_ITEM : n_total;
  int lo[6], hi[6];
  for (int a = 0; a < 6; a++) {
    lo[a] = 0x7fffffff;
    hi[a] = -0x7fffffff;
  }
  for (int64_t p = i0 + threadIdx.x; p < i1; p += 256) {
    const uint32_t c = cell[p];
    if (c == DEAD_CELL) continue;
    const int v[3] = {cell_i(c), cell_j(c), cell_k(c)};
    for (int a = 0; a < 3; a++) {
      const int s = v[a] + nc[a] / 2 < nc[a] ? v[a] + nc[a] / 2 : v[a] + nc[a] / 2 - nc[a];
      lo[a] = min(lo[a], v[a]);
      hi[a] = max(hi[a], v[a]);
      lo[3 + a] = min(lo[3 + a], s);
      hi[3 + a] = max(hi[3 + a], s);
    }
  }
  for (int a = 0; a < 6; a++) {
    const int m0 = wave_min(lo[a]), m1 = wave_max(hi[a]);
    if (__lane_id() == 0) {
      atomicMin(&red[(a / 3) * 6 + a % 3], m0);
      atomicMax(&red[(a / 3) * 6 + 3 + a % 3], m1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    int l = red[a], h = red[3 + a];
    const int ls = red[6 + a], hs = red[9 + a];
    const bool periodic = !(a == 0 && g.ox);
    if (periodic && l <= h && hs - ls < h - l) {  // the shifted frame is tighter
      l = ls - nc[a] / 2;
      h = hs - nc[a] / 2;
    }
    tboxes[6 * (int64_t)blockIdx.x + a] = l;
    tboxes[6 * (int64_t)blockIdx.x + 3 + a] = h;
  }
}

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

}  // namespace


// reset the box/velocity-bound outputs of P3
int pl_box_reset(swap_ctx* c) {
  int64_t n6 = std::max<int64_t>(c->max_items * 6, vb_ints(c));
  hipLaunchKernelGGL(k_pass_reset, dim3(nblk(n6, 256)), dim3(256), 0, c->stream, c->max_items,
                     vb_ints(c), c->boxes_next, c->vb_next, c->dcnt + 4);
  return post(c);
}

int pl_pass(swap_ctx* c, int mode) {
  if (c->n == 0) return 0;
  const int64_t ns = c->n_sorted < 0 ? 0 : c->n_sorted;
  const int64_t tail = c->n - ns, nt = (tail + TAIL_ITEM - 1) / TAIL_ITEM;
  const int64_t nb = (c->n_sorted < 0 ? 0 : c->max_items) + nt;
  if (nb == 0) return 0;
  // the tail's boxes from the cells as they are now (a read of 4 bytes per
  // tail ion); with nothing sorted, every particle is in the tail
  if (nt > 0) {
    hipLaunchKernelGGL(k_tail_boxes, dim3((unsigned)nt), dim3(256), 0, c->stream, ns, c->n, c->cell,
                       c->g, c->tboxes);
    if (int e = post(c)) return e;
  }
  const dim3 grid((unsigned)nb), blk(PASS_WG);
  const int* dn = c->n_sorted < 0 ? c->d_zero : c->d_nitems;
  auto launch = [&](auto R0, auto OX0) {
    using R = decltype(R0);
    constexpr bool OX = decltype(OX0)::value;
    PDev<R> P = pdev<R>(c);
    PCoef<R> K = pcoef<R>(c);
    const R* S = (const R*)c->S;
    using A = typename AccOf<R>::T;
    A* acc = (A*)c->macc;
#define SWAP_PASS(MODE, ACC)                                                                   \
  hipLaunchKernelGGL((k_pass<R, MODE, OX>), grid, blk, 0, c->stream, c->g, P, K, S, ACC,      \
                     c->items, dn, ns, c->n, c->boxes, c->boxes_next, c->vb, c->vb_next,     \
                     c->dcnt, c->tboxes)
    if (mode == 1) SWAP_PASS(1, acc);
    else if (mode == 2) SWAP_PASS(2, acc + c->g.nn);  // up only: moments 1..3
    else SWAP_PASS(3, acc);
#undef SWAP_PASS
  };
  using Open = std::true_type;
  using Periodic = std::false_type;
  if (c->prec == 64) c->g.ox ? launch(0.0, Open()) : launch(0.0, Periodic());
  else c->g.ox ? launch(0.0f, Open()) : launch(0.0f, Periodic());
  return post(c);
}
