// k_denoise.hip -- learned moment denoiser: a 3-D convolutional net trained
// and run on the device with the library's own kernels.
//
// Net: 4 -> 32 -> 32 -> 32 -> 4 channels, 3x3x3 kernels with periodic
// padding, leaky ReLU (slope 1/16 below zero, so no unit can die) after the
// first three layers, residual output (out = in +
// conv4(...)).  Activations are channels-last BF16 [voxel][channel] over the
// interior nodes (i fastest).  Every convolution is an implicit GEMM on MFMA
// (FP32 accumulation): M = voxels, N = output channels, K = 27 taps x input
// channels.  An M-tile is 16 voxels along x; lane l supplies voxel l%16 of
// the A operand and output channel l%16 of the B operand, and receives
// outputs (voxel 4(l/16)+e, channel l%16).  With 32 input channels one
// v_mfma_f32_16x16x32_bf16 takes a whole tap (lane l: channels
// 8(l/16)..8(l/16)+7); with 4, v_mfma_f32_16x16x16_bf16 takes four taps'
// channels in K entries 4(l/16)..4(l/16)+3.
//
// Backward: the loss gradient is propagated with the same kernel through the
// tap-flipped, transposed weights (the adjoint of a periodic convolution),
// scaled by the leaky ReLU's slope of the layer below (1 or 1/16, by the sign
// of its stored activation); weight gradients are a second GEMM
// (M = K entries, N = output channels, reduction over voxels, LDS-staged
// voxel blocks) accumulated into FP32 with atomics.  FP32 master weights,
// Adam.
#include <algorithm>
#include <utility>

#include "swap_internal.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
using bf16 = uint16_t;

// FP32 -> BF16, round to nearest even (finite values)
__device__ __forceinline__ bf16 f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int NL = 4;
constexpr int CH[NL + 1] = {4, 32, 32, 32, 4};
constexpr int64_t wcount(int l) { return 27LL * CH[l] * CH[l + 1]; }
constexpr int64_t layer_params(int l) { return wcount(l) + CH[l + 1]; }
constexpr int64_t param_off(int l) { return l == 0 ? 0 : param_off(l - 1) + layer_params(l - 1); }
constexpr int64_t NPARAM = param_off(NL);
constexpr int k16(int cin) { return (27 * cin + 15) / 16 * 16; }
constexpr int r16(int c) { return (c + 15) / 16 * 16; }

enum { EPI_RELU = 0, EPI_RESID = 1, EPI_MASK = 2 };
constexpr float LEAK = 0.0625f;  // leaky ReLU slope below zero

struct G3 {
  int nx, ny, nz;
  int64_t V;
};

// Y = epilogue(conv(X, W)); Wt is [r16(COUT)][k16(CIN)], zero padded.
// A workgroup computes a CTX x CTY x CTZ block of output voxels.  Its input
// halo (one voxel on each side, periodic, all channels) is staged in LDS
// once; wavefront w runs the four 16-voxel x-runs at y = y0 + w (one per z
// plane) against all output channels, k ascending (tap-major, then channel).
constexpr int CTX = 16, CTY = 4, CTZ = 4;
constexpr int CHX = CTX + 2, CHY = CTY + 2, CHZ = CTZ + 2, CHV = CHX * CHY * CHZ;

unsigned ceil_div(int a, int b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ int wrapi(int s, int n) {
  int r = s % n;
  return r < 0 ? r + n : r;
}

template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv_t(G3 g, const bf16* __restrict__ X,
                                                const bf16* __restrict__ Wt,
                                                const float* __restrict__ bias,
                                                const bf16* __restrict__ mask,
                                                const float* __restrict__ resid,
                                                bf16* __restrict__ Yb, float* __restrict__ Yf) {
  static_assert(CIN == 4 || CIN == 32, "input channels");
  // LDS stride per voxel in BF16: 32 channels padded to 40 (80 B: 16-byte
  // aligned rows, and the 16 voxels of one lane group's 16-byte reads fall
  // in distinct banks)
  constexpr int VS = CIN == 32 ? 40 : CIN;
  constexpr int K = 27 * CIN, KP = k16(CIN), NT = r16(COUT) / 16, CPV = CIN / 4;
  __shared__ __align__(16) bf16 sx[CHV * VS];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int x0 = blockIdx.x * CTX, y0 = blockIdx.y * CTY, z0 = blockIdx.z * CTZ;
  for (int q = tid; q < CHV * CPV; q += 256) {
    const int hv = q / CPV, ch = q % CPV;
    const int hx = hv % CHX, hy = (hv / CHX) % CHY, hz = hv / (CHX * CHY);
    const int64_t gv = wrapi(x0 - 1 + hx, g.nx) +
                       (int64_t)g.nx * (wrapi(y0 - 1 + hy, g.ny) +
                                        (int64_t)g.ny * wrapi(z0 - 1 + hz, g.nz));
    *(s4*)(sx + hv * VS + 4 * ch) = *(const s4*)(X + gv * CIN + 4 * ch);
  }
  __syncthreads();
  const int m = lane & 15, gq = lane >> 4;
  f4 acc[CTZ][NT];
#pragma unroll
  for (int mt = 0; mt < CTZ; mt++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
  if constexpr (CIN == 32) {
    // one tap's 32 channels per MFMA (16x16x32): lane group gq supplies
    // channels 8 gq .. 8 gq + 7 of A (voxel m + tap offset) and of B (weight
    // row m), the same K slots on both sides
    const int ci = 8 * gq;
#pragma unroll 3
    for (int t = 0; t < 27; t++) {
      const int dx = t % 3, dy = (t / 3) % 3, dz = t / 9;
      bf8 a[CTZ];
#pragma unroll
      for (int mt = 0; mt < CTZ; mt++) {
        const int hv = (m + dx) + CHX * ((w + dy) + CHY * (mt + dz));
        a[mt] = __builtin_bit_cast(bf8, *(const s8*)(sx + hv * VS + ci));
      }
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        const bf8 b =
            __builtin_bit_cast(bf8, *(const s8*)(Wt + (int64_t)(nt * 16 + m) * KP + 32 * t + ci));
#pragma unroll
        for (int mt = 0; mt < CTZ; mt++)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b, acc[mt][nt], 0, 0, 0);
      }
    }
  } else
  for (int kk = 0; kk < KP; kk += 16) {
    const int k0 = kk + 4 * gq;
    const bool kok = k0 < K;  // the zero-padded tail of K (taps past 26)
    const int t = !kok ? 0 : (CIN == 32 ? kk / 32 : k0 / 4);
    const int ci = CIN == 32 ? kk % 32 + 4 * gq : 0;
    const int dx = t % 3, dy = (t / 3) % 3, dz = t / 9;
    s4 a[CTZ];
#pragma unroll
    for (int mt = 0; mt < CTZ; mt++) {
      const int hv = (m + dx) + CHX * ((w + dy) + CHY * (mt + dz));
      a[mt] = kok ? *(const s4*)(sx + hv * VS + ci) : s4{0, 0, 0, 0};
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const s4 b = *(const s4*)(Wt + (int64_t)(nt * 16 + m) * KP + k0);
#pragma unroll
      for (int mt = 0; mt < CTZ; mt++)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], b, acc[mt][nt], 0, 0, 0);
    }
  }
  const int y = y0 + w;
  if (y >= g.ny) return;
#pragma unroll
  for (int mt = 0; mt < CTZ; mt++) {
    const int z = z0 + mt;
    if (z >= g.nz) continue;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const int co = nt * 16 + m;
      if (co >= COUT) continue;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int x = x0 + 4 * gq + e;
        if (x >= g.nx) continue;
        const int64_t o = (x + (int64_t)g.nx * (y + (int64_t)g.ny * z)) * COUT + co;
        float val = acc[mt][nt][e];
        if constexpr (EPI == EPI_RELU) {
          val += bias[co];
          Yb[o] = f2bf(val > 0.f ? val : LEAK * val);
        } else if constexpr (EPI == EPI_RESID) {
          Yf[o] = val + bias[co] + resid[o];
        } else {
          Yb[o] = f2bf(bf2f(mask[o]) > 0.f ? val : LEAK * val);
        }
      }
    }
  }
}

// dW[k][co] += sum over voxels v of X[v + offset(tap(k))][ci(k)] dY[v][co], with
// k = tap CIN + ci: an implicit GEMM on MFMA with M = the 27 CIN weight rows
// (tiles of 16 rows: one tap's 16 channels, or 4 taps x 4 channels), N =
// output channels, reduced over voxels in 16-voxel x-runs.  A workgroup
// takes CTX x CTY x CTZ voxel blocks in turn (grid-stride), stages each
// block's input halo (as k_conv_t) and its output gradients in LDS
// (channels-last, 32 channels padded to 36 so the 16 voxels of a lane group
// fall in distinct banks), and keeps its partial dW in registers across the
// blocks: wavefront w owns row tiles w, w + 4, ...; lane l supplies weight
// row l % 16 (A) and output channel l % 16 (B) at the run's voxels
// 4(l/16) .. 4(l/16)+3, and holds rows 4(l/16)+e of channel l % 16.  The
// bias gradient db[co] = sum_v dY[v][co] is summed from the same LDS tiles.
// Each workgroup adds its partials to dW and db once, with FP32 atomics.
constexpr int WG_BLOCKS = 512;  // workgroups of a weight-gradient launch
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(G3 g, const bf16* __restrict__ X,
                                              const bf16* __restrict__ dY, float* dW,
                                              float* db) {
  static_assert(CIN == 4 || CIN == 32, "input channels");
  static_assert(COUT == 4 || COUT == 32, "output channels");
  constexpr int VS = CIN == 32 ? 36 : CIN, YS = COUT == 32 ? 36 : COUT;
  constexpr int K = 27 * CIN, MT = (K + 15) / 16, NT = r16(COUT) / 16, CPV = CIN / 4;
  constexpr int MTW = (MT + 3) / 4, BV = CTX * CTY * CTZ, YPV = COUT / 4;
  __shared__ __align__(16) bf16 sx[CHV * VS];
  __shared__ __align__(16) bf16 sy[BV * YS];
  __shared__ float sb[256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  float bsum = 0.f;  // db partial of channel tid % COUT over voxels tid / COUT + k 256 / COUT
  const int m = lane & 15, gq = lane >> 4;
  const int nbx = (g.nx + CTX - 1) / CTX, nby = (g.ny + CTY - 1) / CTY,
            nbz = (g.nz + CTZ - 1) / CTZ;
  const int nb = nbx * nby * nbz;
  f4 acc[MTW][NT];
#pragma unroll
  for (int j = 0; j < MTW; j++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[j][nt] = f4{0.f, 0.f, 0.f, 0.f};
  for (int blk = blockIdx.x; blk < nb; blk += gridDim.x) {
    const int x0 = (blk % nbx) * CTX, y0 = ((blk / nbx) % nby) * CTY, z0 = blk / (nbx * nby) * CTZ;
    __syncthreads();  // the previous block's LDS reads are done
    for (int q = tid; q < CHV * CPV; q += 256) {
      const int hv = q / CPV, ch = q % CPV;
      const int hx = hv % CHX, hy = (hv / CHX) % CHY, hz = hv / (CHX * CHY);
      const int64_t gv = wrapi(x0 - 1 + hx, g.nx) +
                         (int64_t)g.nx * (wrapi(y0 - 1 + hy, g.ny) +
                                          (int64_t)g.ny * wrapi(z0 - 1 + hz, g.nz));
      *(s4*)(sx + hv * VS + 4 * ch) = *(const s4*)(X + gv * CIN + 4 * ch);
    }
    // output gradients of the block's voxels; voxels past the grid's end add nothing
    for (int q = tid; q < BV * YPV; q += 256) {
      const int lv = q / YPV, ch = q % YPV;
      const int x = x0 + lv % CTX, y = y0 + (lv / CTX) % CTY, z = z0 + lv / (CTX * CTY);
      s4 v = s4{0, 0, 0, 0};
      if (x < g.nx && y < g.ny && z < g.nz)
        v = *(const s4*)(dY + (x + (int64_t)g.nx * (y + (int64_t)g.ny * z)) * COUT + 4 * ch);
      *(s4*)(sy + lv * YS + 4 * ch) = v;
    }
    __syncthreads();
    for (int lv = tid / COUT; lv < BV; lv += 256 / COUT) bsum += bf2f(sy[lv * YS + tid % COUT]);
    for (int r = 0; r < CTY * CTZ; r++) {  // the block's 16-voxel x-runs
      const int ly = r % CTY, lz = r / CTY;
      s4 b[NT];
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        const int co = nt * 16 + m;
#pragma unroll
        for (int e = 0; e < 4; e++)
          b[nt][e] = co < COUT ? (short)sy[(4 * gq + e + CTX * (ly + CTY * lz)) * YS + co] : 0;
      }
#pragma unroll
      for (int j = 0; j < MTW; j++) {
        const int mt = w + 4 * j;
        if (mt >= MT) break;
        const int k = mt * 16 + m;
        s4 a = s4{0, 0, 0, 0};
        if (k < K) {
          const int t = k / CIN, ci = k % CIN;
          const int dx = t % 3, dy = (t / 3) % 3, dz = t / 9;
          const int hb = dx + CHX * ((ly + dy) + CHY * (lz + dz));
#pragma unroll
          for (int e = 0; e < 4; e++) a[e] = (short)sx[(hb + 4 * gq + e) * VS + ci];
        }
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
          acc[j][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], acc[j][nt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < MTW; j++) {
    const int mt = w + 4 * j;
    if (mt >= MT) break;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const int co = nt * 16 + m;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int k = mt * 16 + 4 * gq + e;
        if (k < K && co < COUT) atomicAdd(dW + (int64_t)k * COUT + co, acc[j][nt][e]);
      }
    }
  }
  sb[tid] = bsum;
  __syncthreads();
  if (tid < COUT) {
    float t = 0.f;
    for (int q = tid; q < 256; q += COUT) t += sb[q];
    atomicAdd(db + tid, t);
  }
}

// forward and backward weight layouts from the FP32 master W[tap][ci][co]:
// Wt[co][tap*CIN + ci] (forward), Wb[ci][tap'*COUT + co] = W[26 - tap'][ci][co]
// (the adjoint: input gradient = conv of the output gradient)
template <int CIN, int COUT>
__global__ void k_pack(const float* __restrict__ W, bf16* Wt, bf16* Wb) {
  constexpr int KF = k16(CIN), KB = k16(COUT);
  constexpr int64_t NF = (int64_t)r16(COUT) * KF, NB = (int64_t)r16(CIN) * KB;
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < NF) {
    int co = (int)(q / KF), k = (int)(q % KF);
    float w = 0.f;
    if (co < COUT && k < 27 * CIN) w = W[((int64_t)(k / CIN) * CIN + k % CIN) * COUT + co];
    Wt[q] = f2bf(w);
  } else if (q < NF + NB) {
    int64_t r = q - NF;
    int ci = (int)(r / KB), k = (int)(r % KB);
    float w = 0.f;
    if (ci < CIN && k < 27 * COUT) {
      int t = k / COUT, co = k % COUT;
      w = W[((int64_t)(26 - t) * CIN + ci) * COUT + co];
    }
    Wb[r] = f2bf(w);
  }
}

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// mean squared error over n values; gradient 2 (out - tgt) / n as BF16
__global__ void k_mse(int64_t n, const float* __restrict__ out, const float* __restrict__ tgt,
                      bf16* __restrict__ grad, double* loss) {
  double s = 0.0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    float d = out[q] - tgt[q];
    if (grad) grad[q] = f2bf(2.f * d / (float)n);
    s += (double)d * d;
  }
  s = wsum(s);
  if (__lane_id() == 0) atomicAdd(loss, s / (double)n);
}

__global__ void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr,
                       float c1, float c2) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  float gq = g[q];
  float mq = 0.9f * m[q] + 0.1f * gq;
  float vq = 0.999f * v[q] + 0.001f * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
}

// io = (S A - io) / (S - 1) per moment sum: fixed-point sums (two's
// complement in 64 bits, rounded back to the grid) or doubles
template <typename T>
__global__ void k_rest(int64_t n, const T* __restrict__ A, T* __restrict__ io, int S) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  if constexpr (sizeof(T) == 8 && !__is_same(T, double)) {
    const long long dlt = (long long)S * (long long)A[q] - (long long)io[q];
    io[q] = (T)(long long)llrint((double)dlt / (double)(S - 1));
  } else {
    io[q] = ((T)S * A[q] - io[q]) / (T)(S - 1);
  }
}

__global__ void k_f2bf(int64_t n, const float* __restrict__ x, bf16* __restrict__ y) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) y[q] = f2bf(x[q]);
}

// He-normal weights (Box-Muller on Philox draws), zero biases; the last
// layer starts small, so the residual net starts near the identity
__global__ void k_init(float* p, uint32_t k0, uint32_t k1) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  int l = 0;
  while (l + 1 < NL && q >= param_off(l + 1)) l++;
  int64_t r = q - param_off(l);
  if (r >= wcount(l)) {
    p[q] = 0.f;
    return;
  }
  uint32_t c[4] = {(uint32_t)q, 0u, 0x50000u, 0u};
  philox4x32_10(c, k0, k1);
  double u1 = 1.0 - u24(c[0]), u2 = u24(c[1]);
  double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  double sd = sqrt(2.0 / (27.0 * CH[l])) * (l == NL - 1 ? 0.1 : 1.0);
  p[q] = (float)(z * sd);
}

// moments (storage layout, ghosts included) -> normalised [voxel][4] FP32
template <typename R>
__global__ void k_dn_in(Geo g, const R* __restrict__ np, const R* __restrict__ up, float nsc,
                        float vsc, float* __restrict__ X) {
  int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t V = (int64_t)g.ncx * g.ncy * g.ncz;
  if (v >= V) return;
  int i = 1 + (int)(v % g.ncx), j = 1 + (int)((v / g.ncx) % g.ncy),
      k = 1 + (int)(v / ((int64_t)g.ncx * g.ncy));
  int64_t s = gidx(g, i, j, k);
  X[4 * v] = (float)np[s] / nsc - 1.f;
  for (int c = 0; c < 3; c++) X[4 * v + 1 + c] = (float)up[s + c * g.nn] / vsc;
}

// net output -> moments at every storage node (ghosts = periodic images)
template <typename R>
__global__ void k_dn_out(Geo g, const float* __restrict__ Y, float nsc, float vsc, int with_np,
                         R* np, R* up) {
  int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= g.nn) return;
  int i = (int)(s % g.nx), j = (int)((s / g.nx) % g.ny), k = (int)(s / ((int64_t)g.nx * g.ny));
  int64_t v = (int64_t)(canon(i, g.nx) - 1) +
              (int64_t)g.ncx * ((int64_t)(canon(j, g.ny) - 1) + (int64_t)g.ncy * (canon(k, g.nz) - 1));
  if (with_np) np[s] = (R)((Y[4 * v] + 1.f) * nsc);
  for (int c = 0; c < 3; c++) up[s + c * g.nn] = (R)(Y[4 * v + 1 + c] * vsc);
}

int post(swap_ctx* c) {
  HIPCHK(c, hipGetLastError());
  return 0;
}

}  // namespace

struct DnState {
  swap_dn_config cfg;
  G3 g3;
  float* params = nullptr;  // FP32 master weights, flat
  float* grads = nullptr;
  float* am = nullptr;      // Adam moments
  float* av = nullptr;
  bf16* Wt[NL] = {};
  bf16* Wb[NL] = {};
  bf16* act[NL] = {};       // act[0] input (4 ch), act[1..3] after the leaky ReLU (32 ch)
  bf16* gr[NL + 1] = {};    // gradients w.r.t. the inputs of layer l (pre-activation of l-1); gr[NL] = dOut
  float* xf = nullptr;      // normalised input [V][4]
  float* out = nullptr;     // output [V][4]
  double* dloss = nullptr;
  float* pin = nullptr;     // pairs: inputs [max_pairs][V][4]
  float* ptg = nullptr;     // pairs: targets
  void* sacc = nullptr;     // moment sums of a subsample [4][nn] (stride > 1)
  int npairs = 0;
  int64_t t = 0;            // Adam step
  bool gate = false, enabled = false;
};

namespace {

int dn_pack(swap_ctx* c) {
  DnState* d = c->dn;
  auto pk = [&](auto kern, int l, int64_t nf, int64_t nb) {
    hipLaunchKernelGGL(kern, dim3(nblk(nf + nb, 256)), dim3(256), 0, c->stream,
                       d->params + param_off(l), d->Wt[l], d->Wb[l]);
  };
  pk(k_pack<4, 32>, 0, (int64_t)r16(32) * k16(4), (int64_t)r16(4) * k16(32));
  pk(k_pack<32, 32>, 1, (int64_t)r16(32) * k16(32), (int64_t)r16(32) * k16(32));
  pk(k_pack<32, 32>, 2, (int64_t)r16(32) * k16(32), (int64_t)r16(32) * k16(32));
  pk(k_pack<32, 4>, 3, (int64_t)r16(4) * k16(32), (int64_t)r16(32) * k16(4));
  return post(c);
}

// forward from d->xf to d->out, keeping the activations
int dn_forward(swap_ctx* c) {
  DnState* d = c->dn;
  const G3 g = d->g3;
  const int64_t V = g.V;
  const dim3 grid(ceil_div(g.nx, CTX), ceil_div(g.ny, CTY), ceil_div(g.nz, CTZ)), blk(256);
  hipLaunchKernelGGL(k_f2bf, dim3(nblk(4 * V, 256)), dim3(256), 0, c->stream, 4 * V, d->xf,
                     d->act[0]);
  const float* b[NL];
  for (int l = 0; l < NL; l++) b[l] = d->params + param_off(l) + wcount(l);
  hipLaunchKernelGGL((k_conv_t<4, 32, EPI_RELU>), grid, blk, 0, c->stream, g, d->act[0], d->Wt[0],
                     b[0], nullptr, nullptr, d->act[1], nullptr);
  hipLaunchKernelGGL((k_conv_t<32, 32, EPI_RELU>), grid, blk, 0, c->stream, g, d->act[1],
                     d->Wt[1], b[1], nullptr, nullptr, d->act[2], nullptr);
  hipLaunchKernelGGL((k_conv_t<32, 32, EPI_RELU>), grid, blk, 0, c->stream, g, d->act[2],
                     d->Wt[2], b[2], nullptr, nullptr, d->act[3], nullptr);
  hipLaunchKernelGGL((k_conv_t<32, 4, EPI_RESID>), grid, blk, 0, c->stream, g, d->act[3],
                     d->Wt[3], b[3], nullptr, d->xf, nullptr, d->out);
  return post(c);
}

// loss against tgt and the parameter gradients into d->grads
int dn_backward(swap_ctx* c, const float* tgt) {
  DnState* d = c->dn;
  const G3 g = d->g3;
  const int64_t V = g.V;
  const dim3 grid(ceil_div(g.nx, CTX), ceil_div(g.ny, CTY), ceil_div(g.nz, CTZ)), blk(256);
  HIPCHK(c, hipMemsetAsync(d->grads, 0, sizeof(float) * NPARAM, c->stream));
  HIPCHK(c, hipMemsetAsync(d->dloss, 0, sizeof(double), c->stream));
  hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, c->stream, 4 * V, d->out, tgt, d->gr[NL],
                     d->dloss);
  const int nvb = (int)(ceil_div(g.nx, CTX) * ceil_div(g.ny, CTY) * ceil_div(g.nz, CTZ));
  const dim3 wgrid((unsigned)std::min(nvb, WG_BLOCKS)), wblk(256);
  float* gw[NL];
  float* gb[NL];
  for (int l = 0; l < NL; l++) {
    gw[l] = d->grads + param_off(l);
    gb[l] = gw[l] + wcount(l);
  }
  // layer 4 (32 -> 4, no activation)
  hipLaunchKernelGGL((k_wgrad<32, 4>), wgrid, wblk, 0, c->stream, g, d->act[3], d->gr[4], gw[3],
                     gb[3]);
  hipLaunchKernelGGL((k_conv_t<4, 32, EPI_MASK>), grid, blk, 0, c->stream, g, d->gr[4], d->Wb[3],
                     nullptr, d->act[3], nullptr, d->gr[3], nullptr);
  // layers 3 and 2 (32 -> 32)
  for (int l = 2; l >= 1; l--) {
    hipLaunchKernelGGL((k_wgrad<32, 32>), wgrid, wblk, 0, c->stream, g, d->act[l], d->gr[l + 1],
                       gw[l], gb[l]);
    hipLaunchKernelGGL((k_conv_t<32, 32, EPI_MASK>), grid, blk, 0, c->stream, g, d->gr[l + 1],
                       d->Wb[l], nullptr, d->act[l], nullptr, d->gr[l], nullptr);
  }
  // layer 1 (4 -> 32); no gradient to the input
  hipLaunchKernelGGL((k_wgrad<4, 32>), wgrid, wblk, 0, c->stream, g, d->act[0], d->gr[1], gw[0],
                     gb[0]);
  return post(c);
}

int dn_moments_in(swap_ctx* c, float* X) {
  DnState* d = c->dn;
  if (c->prec == 64)
    hipLaunchKernelGGL(k_dn_in<double>, dim3(nblk(d->g3.V, 256)), dim3(256), 0, c->stream, c->g,
                       (const double*)c->fld[SWAP_F_NP], (const double*)c->fld[SWAP_F_UP],
                       (float)d->cfg.n_scale, (float)d->cfg.v_scale, X);
  else
    hipLaunchKernelGGL(k_dn_in<float>, dim3(nblk(d->g3.V, 256)), dim3(256), 0, c->stream, c->g,
                       (const float*)c->fld[SWAP_F_NP], (const float*)c->fld[SWAP_F_UP],
                       (float)d->cfg.n_scale, (float)d->cfg.v_scale, X);
  return post(c);
}

int zero_acc(swap_ctx* c, int m0, int m1) {
  c->mgen++;
  size_t off = 8 * (size_t)c->g.nn * m0, len = 8 * (size_t)c->g.nn * (m1 - m0);
  HIPCHK(c, hipMemsetAsync((char*)c->macc + off, 0, len, c->stream));
  return 0;
}

// np and up of the current particles (all, or the stride subsample) into the
// context's np/up fields
int dn_deposit(swap_ctx* c, int stride) {
  if (int e = zero_acc(c, 0, 4)) return e;
  if (int e = pl_deposit_simple(c, 0, stride)) return e;
  if (int e = fl_finalize_np(c)) return e;
  if (int e = pl_deposit_simple(c, 1, stride)) return e;
  return fl_finalize_up(c);
}

int dn_need(swap_ctx* c) {
  if (!c->dn) return set_err(c, SWAP_ESTATE, "denoiser: swap_dn_create first");
  return 0;
}

// [4][V] host -> [V][4] device
int host_to_dev(swap_ctx* c, const float* x, float* dst) {
  const int64_t V = c->dn->g3.V;
  std::vector<float> t(4 * V);
  for (int64_t v = 0; v < V; v++)
    for (int q = 0; q < 4; q++) t[4 * v + q] = x[q * V + v];
  HIPCHK(c, hipMemcpyAsync(dst, t.data(), 16 * V, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// [V][4] device -> [4][V] host
int dev_to_host(swap_ctx* c, const float* src, float* x) {
  const int64_t V = c->dn->g3.V;
  std::vector<float> t(4 * V);
  HIPCHK(c, hipMemcpyAsync(t.data(), src, 16 * V, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t v = 0; v < V; v++)
    for (int q = 0; q < 4; q++) x[q * V + v] = t[4 * v + q];
  return 0;
}

}  // namespace

// ---------------- C-ABI ----------------

extern "C" int64_t swap_dn_num_params(void) { return NPARAM; }

extern "C" int swap_dn_create(swap_ctx* c, const swap_dn_config* cfg) {
  if (!c || !cfg) return SWAP_EINVAL;
  if (!(cfg->n_scale > 0) || !(cfg->v_scale > 0) || cfg->max_pairs < 0 || cfg->stride < 1)
    return set_err(c, SWAP_EINVAL, "dn_create: bad config");
  if (c->g.ox)  // the convolutions pad periodically in every axis
    return set_err(c, SWAP_EINVAL, "dn_create: the denoiser needs a periodic x");
  dn_destroy(c);
  DnState* d = new DnState();
  c->dn = d;
  d->cfg = *cfg;
  d->g3 = G3{c->g.ncx, c->g.ncy, c->g.ncz, (int64_t)c->g.ncx * c->g.ncy * c->g.ncz};
  const int64_t V = d->g3.V;
  auto A = [&](void** p, size_t b) -> int {
    hipError_t e = hipMalloc(p, b ? b : 16);
    if (e != hipSuccess)
      return set_err(c, SWAP_ENOMEM, "dn_create: hipMalloc(%zu): %s", b, hipGetErrorString(e));
    return 0;
  };
#define DA(p, b) \
  if (int e_ = A((void**)&(p), (b))) return e_
  DA(d->params, 4 * NPARAM);
  DA(d->grads, 4 * NPARAM);
  DA(d->am, 4 * NPARAM);
  DA(d->av, 4 * NPARAM);
  for (int l = 0; l < NL; l++) {
    DA(d->Wt[l], 2 * (size_t)r16(CH[l + 1]) * k16(CH[l]));
    DA(d->Wb[l], 2 * (size_t)r16(CH[l]) * k16(CH[l + 1]));
    DA(d->act[l], 2 * (size_t)V * CH[l]);
    DA(d->gr[l + 1], 2 * (size_t)V * CH[l + 1]);
  }
  DA(d->xf, 16 * (size_t)V);
  DA(d->out, 16 * (size_t)V);
  DA(d->dloss, sizeof(double));
  DA(d->pin, 16 * (size_t)V * (size_t)cfg->max_pairs);
  DA(d->ptg, 16 * (size_t)V * (size_t)cfg->max_pairs);
  if (cfg->stride > 1) DA(d->sacc, 32 * (size_t)c->g.nn);
#undef DA
  HIPCHK(c, hipMemsetAsync(d->am, 0, 4 * NPARAM, c->stream));
  HIPCHK(c, hipMemsetAsync(d->av, 0, 4 * NPARAM, c->stream));
  hipLaunchKernelGGL(k_init, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, d->params,
                     (uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32));
  if (int e = post(c)) return e;
  if (int e = dn_pack(c)) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

void dn_destroy(swap_ctx* c) {
  DnState* d = c->dn;
  if (!d) return;
  void* ps[] = {d->params, d->grads, d->am, d->av, d->xf,
                d->out,    d->dloss, d->pin, d->ptg, d->sacc};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (int l = 0; l < NL; l++) {
    void* q[] = {d->Wt[l], d->Wb[l], d->act[l], d->gr[l + 1]};
    for (void* p : q)
      if (p) (void)hipFree(p);
  }
  delete d;
  c->dn = nullptr;
}

// the input and target moments of SWAP_DN_DISJOINT_STEP: macc holds the
// step's sums A over all particles at x^{n+1/2} (before the final half move)
// and stays untouched; the subsample (weights x S) is deposited at its
// x^{n+1/2}, half a move back, into the scratch sums, and the target is
// (S A - scratch) / (S - 1) = the other particles with weights x S/(S-1)
static int dn_pair_from_step(swap_ctx* c, float* in, float* tg) {
  DnState* d = c->dn;
  const int S = d->cfg.stride;
  const int64_t n4 = 4 * c->g.nn;
  struct Swap {  // macc <-> scratch for the finalize kernels, restored on exit
    swap_ctx* c;
    ~Swap() { std::swap(c->macc, c->dn->sacc); }
  };
  std::swap(c->macc, d->sacc);
  Swap guard{c};
  HIPCHK(c, hipMemsetAsync(c->macc, 0, 8 * (size_t)n4, c->stream));
  if (int e = pl_deposit_simple(c, 0, S, true)) return e;
  if (int e = fl_finalize_np(c)) return e;
  if (int e = pl_deposit_simple(c, 1, S, true)) return e;
  if (int e = fl_finalize_up(c)) return e;
  if (int e = dn_moments_in(c, in)) return e;
  if (c->prec == 64)
    hipLaunchKernelGGL(k_rest<double>, dim3(nblk(n4, 256)), dim3(256), 0, c->stream, n4,
                       (const double*)d->sacc, (double*)c->macc, S);
  else
    hipLaunchKernelGGL(k_rest<unsigned long long>, dim3(nblk(n4, 256)), dim3(256), 0, c->stream,
                       n4, (const unsigned long long*)d->sacc, (unsigned long long*)c->macc, S);
  if (int e = post(c)) return e;
  if (int e = fl_finalize_np(c)) return e;
  if (int e = fl_finalize_up(c)) return e;
  return dn_moments_in(c, tg);
}

extern "C" int swap_dn_add_pair(swap_ctx* c, int mode) {
  if (!c || mode < SWAP_DN_ALL || mode > SWAP_DN_DISJOINT_STEP) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  if (d->npairs >= d->cfg.max_pairs) return set_err(c, SWAP_ECAP, "dn_add_pair: pairs full");
  if (!c->id && d->cfg.stride > 1)
    return set_err(c, SWAP_ESTATE, "dn_add_pair needs SWAP_TRACK_IDS");
  if (mode != SWAP_DN_ALL && d->cfg.stride < 2)
    return set_err(c, SWAP_EINVAL, "dn_add_pair: a disjoint target needs stride > 1");
  if (mode == SWAP_DN_DISJOINT_STEP && c->mgen != c->step_gen)
    return set_err(c, SWAP_ESTATE,
                   "dn_add_pair: particles or moment sums changed since the last swap_step");
  const size_t nn = (size_t)c->g.nn, rs = c->rsz;
  // keep the context's np/up
  HIPCHK(c, hipMemcpyAsync(c->fld[SWAP_F_TMP2], c->fld[SWAP_F_NP], rs * nn,
                           hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->fld[SWAP_F_TMP1], c->fld[SWAP_F_UP], 3 * rs * nn,
                           hipMemcpyDeviceToDevice, c->stream));
  const int64_t V = d->g3.V;
  float* in = d->pin + (size_t)d->npairs * 4 * V;
  float* tg = d->ptg + (size_t)d->npairs * 4 * V;
  if (mode == SWAP_DN_DISJOINT_STEP) {
    if (int e = dn_pair_from_step(c, in, tg)) return e;
  } else {
    if (int e = dn_deposit(c, d->cfg.stride)) return e;
    if (int e = dn_moments_in(c, in)) return e;
    if (int e = dn_deposit(c, mode == SWAP_DN_DISJOINT ? -d->cfg.stride : 1)) return e;
    if (int e = dn_moments_in(c, tg)) return e;
  }
  HIPCHK(c, hipMemcpyAsync(c->fld[SWAP_F_NP], c->fld[SWAP_F_TMP2], rs * nn,
                           hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->fld[SWAP_F_UP], c->fld[SWAP_F_TMP1], 3 * rs * nn,
                           hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d->npairs++;
  return 0;
}

extern "C" int swap_dn_num_pairs(swap_ctx* c) { return c && c->dn ? c->dn->npairs : 0; }

extern "C" int swap_dn_get_pair(swap_ctx* c, int k, float* x, float* target) {
  if (!c || !x || !target) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  if (k < 0 || k >= d->npairs) return set_err(c, SWAP_EINVAL, "dn_get_pair: k");
  const int64_t V = d->g3.V;
  if (int e = dev_to_host(c, d->pin + (size_t)k * 4 * V, x)) return e;
  return dev_to_host(c, d->ptg + (size_t)k * 4 * V, target);
}

extern "C" int swap_dn_put_pair(swap_ctx* c, const float* x, const float* target) {
  if (!c || !x || !target) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  if (d->npairs >= d->cfg.max_pairs) return set_err(c, SWAP_ECAP, "dn_put_pair: pairs full");
  const int64_t V = d->g3.V;
  if (int e = host_to_dev(c, x, d->pin + (size_t)d->npairs * 4 * V)) return e;
  if (int e = host_to_dev(c, target, d->ptg + (size_t)d->npairs * 4 * V)) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d->npairs++;
  return 0;
}

extern "C" int swap_dn_train(swap_ctx* c, int iters, double lr, int ntrain, double* loss) {
  if (!c || iters < 0 || !(lr > 0)) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  if (ntrain < 1 || ntrain > d->npairs) return set_err(c, SWAP_EINVAL, "dn_train: ntrain");
  const int64_t V = d->g3.V;
  for (int it = 0; it < iters; it++) {
    uint32_t r[4] = {(uint32_t)d->t, 0u, 0x60000u, 0u};
    philox4x32_10(r, (uint32_t)d->cfg.seed, (uint32_t)(d->cfg.seed >> 32));
    int pidx = (int)(r[0] % (uint32_t)ntrain);
    HIPCHK(c, hipMemcpyAsync(d->xf, d->pin + (size_t)pidx * 4 * V, 16 * V,
                             hipMemcpyDeviceToDevice, c->stream));
    if (int e = dn_forward(c)) return e;
    if (int e = dn_backward(c, d->ptg + (size_t)pidx * 4 * V)) return e;
    d->t++;
    float c1 = (float)(1.0 - pow(0.9, (double)d->t)), c2 = (float)(1.0 - pow(0.999, (double)d->t));
    hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, NPARAM,
                       d->params, d->grads, d->am, d->av, (float)lr, c1, c2);
    if (int e = post(c)) return e;
    if (int e = dn_pack(c)) return e;
  }
  double h = 0.0;
  HIPCHK(c, hipMemcpyAsync(&h, d->dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = h;
  return 0;
}

extern "C" int swap_dn_eval(swap_ctx* c, int p0, int p1, double* err_raw, double* err_net) {
  if (!c || !err_raw || !err_net) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  if (p0 < 0 || p1 > d->npairs || p0 >= p1) return set_err(c, SWAP_EINVAL, "dn_eval: range");
  const int64_t V = d->g3.V;
  double sr = 0.0, sn = 0.0;
  for (int p = p0; p < p1; p++) {
    const float* in = d->pin + (size_t)p * 4 * V;
    const float* tg = d->ptg + (size_t)p * 4 * V;
    double h;
    HIPCHK(c, hipMemsetAsync(d->dloss, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, c->stream, 4 * V, in, tg, nullptr,
                       d->dloss);
    HIPCHK(c, hipMemcpyAsync(&h, d->dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    sr += h;
    HIPCHK(c, hipMemcpyAsync(d->xf, in, 16 * V, hipMemcpyDeviceToDevice, c->stream));
    if (int e = dn_forward(c)) return e;
    HIPCHK(c, hipMemsetAsync(d->dloss, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, c->stream, 4 * V, d->out, tg, nullptr,
                       d->dloss);
    HIPCHK(c, hipMemcpyAsync(&h, d->dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    sn += h;
  }
  *err_raw = sr / (p1 - p0);
  *err_net = sn / (p1 - p0);
  d->gate = *err_net < *err_raw;
  if (!d->gate) d->enabled = false;
  return 0;
}

extern "C" int swap_dn_enable(swap_ctx* c, int on) {
  if (!c) return SWAP_EINVAL;
  if (!on) {
    if (c->dn) c->dn->enabled = false;
    return 0;
  }
  if (int e = dn_need(c)) return e;
  if (!c->dn->gate)
    return set_err(c, SWAP_ESTATE,
                   "denoiser gate closed: the last swap_dn_eval did not beat plain deposition");
  c->dn->enabled = true;
  return 0;
}

extern "C" int swap_dn_grad(swap_ctx* c, const float* x, const float* target, double* loss,
                            float* grads) {
  if (!c || !x || !target || !loss || !grads) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  const int64_t V = d->g3.V;
  float* tdev = nullptr;
  HIPCHK(c, hipMalloc(&tdev, 16 * V));
  int e = host_to_dev(c, target, tdev);
  if (!e) e = host_to_dev(c, x, d->xf);
  if (!e) e = dn_forward(c);
  if (!e) e = dn_backward(c, tdev);
  if (!e) {
    HIPCHK(c, hipMemcpyAsync(grads, d->grads, 4 * NPARAM, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(loss, d->dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  (void)hipFree(tdev);
  return e;
}

extern "C" int swap_dn_infer(swap_ctx* c, const float* x, float* y) {
  if (!c || !x || !y) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  DnState* d = c->dn;
  const int64_t V = d->g3.V;
  if (int e = host_to_dev(c, x, d->xf)) return e;
  if (int e = dn_forward(c)) return e;
  std::vector<float> t(4 * V);
  HIPCHK(c, hipMemcpyAsync(t.data(), d->out, 16 * V, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t v = 0; v < V; v++)
    for (int q = 0; q < 4; q++) y[q * V + v] = t[4 * v + q];
  return 0;
}

extern "C" int swap_dn_get_params(swap_ctx* c, float* params) {
  if (!c || !params) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  HIPCHK(c, hipMemcpyAsync(params, c->dn->params, 4 * NPARAM, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int swap_dn_set_params(swap_ctx* c, const float* params) {
  if (!c || !params) return SWAP_EINVAL;
  if (int e = dn_need(c)) return e;
  HIPCHK(c, hipMemcpyAsync(c->dn->params, params, 4 * NPARAM, hipMemcpyHostToDevice, c->stream));
  if (int e = dn_pack(c)) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int dn_apply_step(swap_ctx* c, bool with_np) {
  DnState* d = c->dn;
  if (!d || !d->enabled) return 0;
  if (int e = dn_moments_in(c, d->xf)) return e;
  if (int e = dn_forward(c)) return e;
  if (c->prec == 64)
    hipLaunchKernelGGL(k_dn_out<double>, dim3(nblk(c->g.nn, 256)), dim3(256), 0, c->stream,
                       c->g, d->out, (float)d->cfg.n_scale, (float)d->cfg.v_scale,
                       with_np ? 1 : 0, (double*)c->fld[SWAP_F_NP], (double*)c->fld[SWAP_F_UP]);
  else
    hipLaunchKernelGGL(k_dn_out<float>, dim3(nblk(c->g.nn, 256)), dim3(256), 0, c->stream,
                       c->g, d->out, (float)d->cfg.n_scale, (float)d->cfg.v_scale,
                       with_np ? 1 : 0, (float*)c->fld[SWAP_F_NP], (float*)c->fld[SWAP_F_UP]);
  return post(c);
}
