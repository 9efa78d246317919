// Learned surrogate: an ensemble of MLPs 10 -> 256 -> 256 -> 256 -> 6 (SiLU)
// trained and evaluated on the gfx950 matrix cores.  Weights and
// activations are BF16 (v_mfma_f32_16x16x32_bf16, FP32 accumulate); biases,
// FP32 master weights and Adam moments stay FP32.
//
// MFMA 16x16x32 operand layout (wave64): lane l supplies row (l & 15) of A
// and column (l & 15) of B, each as the 8 consecutive k values
// 8*(l >> 4) .. 8*(l >> 4) + 7; the 16x16 FP32 result leaves lane l holding
// rows 4*(l >> 4) .. 4*(l >> 4) + 3 of column (l & 15).  Every operand here
// is stored with k contiguous, so each lane's fragment is one 16-byte load.
//
// Training (one launch per layer and pass, all members in grid.z):
//   prep       gather the mini-batch, normalise, round to BF16
//   fwd 1..4   Z = A W^T + b; hidden layers store H = silu(Z), H^T and
//              silu'(Z) (BF16); the output layer stores Y (FP32)
//   loss       dL/dY of the mean squared error (BF16, both layouts)
//   bwd 4..2   D = (D_next W) * silu'(Z)  (W^T copies as the B operand)
//   dW 1..4    D^T H over the batch, split over K into FP32 partials; the
//              transposed activations carry a row of ones, so the bias
//              gradient is one more column of the same product
//   adam       sums the partials in a fixed order (deterministic), updates
//              the FP32 masters and rewrites the BF16 copies
// Inference: one fused kernel per tile of samples keeps the activations in
// LDS through all four layers and all members; with a Jacobian it carries six
// tangent planes (d/d e_j, forward mode) through the same MFMA products.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dev_util.hpp"
#include "mf_internal.hpp"
#include "mlp_internal.hpp"

namespace mf {
namespace mlpk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// BF16 copies of the weights, per member (element counts).
constexpr int kInP = 32;     // input width padded to one MFMA k step
constexpr int kOutP = 16;    // output width padded to one MFMA tile
constexpr int kTRows = 272;  // transposed hidden activations: 256 + ones row + pad
constexpr int64_t kW1 = kHid * kInP, kW2 = kHid * kHid, kW4 = kOutP * kHid;
constexpr int64_t kW4T = kHid * kInP;
constexpr int64_t kBias = 3 * kHid + kOutP;
constexpr int kActLd = kHid + 8;  // LDS row stride (BF16) of the inference tile

__device__ __forceinline__ uint16_t bf(float x) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x));
}
__device__ __forceinline__ float unbf(uint16_t b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
// two floats to BF16 (round to nearest even) in one 32-bit word, low half
// first: one v_cvt_pk_bf16_f32 (converted one at a time, each value took a
// conversion and the pair a shift and an or)
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
  return make_uint2(pack2(a, b), pack2(c, d));
}
// silu(z) = z * sigmoid(z); ds = silu'(z) = sg * (1 + z * (1 - sg))
__device__ __forceinline__ float silu(float z, float &ds) {
  const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  ds = sg * (1.0f + z * (1.0f - sg));
  return z * sg;
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// ---------------------------------------------------------------------------
// Generic C = A B^T (A: M x K, B: N x K, both BF16 with K contiguous).  Each
// wave computes one 32 x 32 tile (2 x 2 MFMA tiles); grid.y splits K,
// grid.z is the ensemble member.  M and N are multiples of 16, kc of 32.
enum Epi { kFwdHidden = 0, kFwdOut = 1, kBwdHidden = 2, kPartial = 3 };

struct GemmArgs {
  const uint16_t *A, *B;
  int64_t lda, ldb, sA, sB;
  int M, N, kc;
  const float *bias;
  int64_t sbias;
  const uint16_t *Sin;  // silu'(Z) of the layer below (kBwdHidden), ld ldo
  uint16_t *O, *OT, *OS;
  int64_t ldo, ldt, sO, sOT;
  float *Y;
  int64_t ldy, sY, sYsplit;
};

template <int EPI>
__device__ __forceinline__ void gemm_store(const GemmArgs &g, f32x4 v, int rb, int cb,
                                           int mem, int split, int lane) {
  const int col = cb + (lane & 15);
  const int row = rb + (lane >> 4) * 4;  // rows row .. row + 3
  if constexpr (EPI == kPartial) {
    float *Y = g.Y + mem * g.sY + split * g.sYsplit + col;
#pragma unroll
    for (int q = 0; q < 4; ++q) Y[(int64_t)(row + q) * g.ldy] = v[q];
  } else if constexpr (EPI == kFwdOut) {
    const float b = g.bias[mem * g.sbias + col];
    float *Y = g.Y + mem * g.sY + col;
#pragma unroll
    for (int q = 0; q < 4; ++q) Y[(int64_t)(row + q) * g.ldy] = v[q] + b;
  } else if constexpr (EPI == kFwdHidden) {
    const float b = g.bias[mem * g.sbias + col];
    float h[4], d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = silu(v[q] + b, d[q]);
    uint16_t *O = g.O + mem * g.sO + col, *S = g.OS + mem * g.sO + col;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      O[(int64_t)(row + q) * g.ldo] = bf(h[q]);
      S[(int64_t)(row + q) * g.ldo] = bf(d[q]);
    }
    *reinterpret_cast<uint2 *>(g.OT + mem * g.sOT + (int64_t)col * g.ldt + row) =
        pack4(h[0], h[1], h[2], h[3]);
  } else {  // kBwdHidden
    const uint16_t *S = g.Sin + mem * g.sO + col;
    float d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = v[q] * unbf(S[(int64_t)(row + q) * g.ldo]);
    uint16_t *O = g.O + mem * g.sO + col;
#pragma unroll
    for (int q = 0; q < 4; ++q) O[(int64_t)(row + q) * g.ldo] = bf(d[q]);
    *reinterpret_cast<uint2 *>(g.OT + mem * g.sOT + (int64_t)col * g.ldt + row) =
        pack4(d[0], d[1], d[2], d[3]);
  }
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int tiles_n = (g.N + 31) >> 5;
  const int tiles = ((g.M + 31) >> 5) * tiles_n;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= tiles) return;  // whole wave; the kernel has no barriers
  const int r0 = (tile / tiles_n) * 32, c0 = (tile % tiles_n) * 32;
  const bool row1 = r0 + 16 < g.M, col1 = c0 + 16 < g.N;
  const int mem = blockIdx.z, split = blockIdx.y;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const int64_t k0 = (int64_t)split * g.kc + lk;
  const uint16_t *pa0 = g.A + mem * g.sA + (int64_t)(r0 + lr) * g.lda + k0;
  const uint16_t *pa1 = row1 ? pa0 + 16 * g.lda : pa0;
  const uint16_t *pb0 = g.B + mem * g.sB + (int64_t)(c0 + lr) * g.ldb + k0;
  const uint16_t *pb1 = col1 ? pb0 + 16 * g.ldb : pb0;
  f32x4 a00 = {0.f, 0.f, 0.f, 0.f}, a01 = a00, a10 = a00, a11 = a00;
  for (int k = 0; k < g.kc; k += 32) {
    const bf16x8 x0 = ld8(pa0 + k), x1 = ld8(pa1 + k);
    const bf16x8 y0 = ld8(pb0 + k), y1 = ld8(pb1 + k);
    a00 = mfma(x0, y0, a00);
    a01 = mfma(x0, y1, a01);
    a10 = mfma(x1, y0, a10);
    a11 = mfma(x1, y1, a11);
  }
  gemm_store<EPI>(g, a00, r0, c0, mem, split, lane);
  if (col1) gemm_store<EPI>(g, a01, r0, c0 + 16, mem, split, lane);
  if (row1) {
    gemm_store<EPI>(g, a10, r0 + 16, c0, mem, split, lane);
    if (col1) gemm_store<EPI>(g, a11, r0 + 16, c0 + 16, mem, split, lane);
  }
}

// Gathers member k's mini-batch rows idx[k*idx_stride + b], normalises them
// and writes the BF16 inputs in both layouts plus the FP32 targets.
__global__ void prep_kernel(const float *X, const float *Y, int64_t n_data,
                            const int32_t *idx, int64_t idx_stride, int Bt,
                            const MlpNorm *nm, uint16_t *Xn, uint16_t *XnT,
                            float *Yt) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int mem = blockIdx.y;
  if (b >= Bt) return;
  int64_t r = idx ? idx[mem * idx_stride + b] : b;
  r = r < 0 ? 0 : (r >= n_data ? n_data - 1 : r);
  uint16_t *xn = Xn + (int64_t)mem * Bt * kInP + (int64_t)b * kInP;
  uint16_t *xt = XnT + (int64_t)mem * kInP * Bt + b;
  for (int i = 0; i < kIn; ++i) {
    const uint16_t u = bf((X[r * kIn + i] - nm->in_mean[i]) * nm->in_inv[i]);
    xn[i] = u;
    xt[(int64_t)i * Bt] = u;
  }
  float *yt = Yt + (int64_t)mem * Bt * 8 + (int64_t)b * 8;
  for (int i = 0; i < kOut; ++i) yt[i] = (Y[r * kOut + i] - nm->out_mean[i]) * nm->out_inv[i];
}

__device__ float block_sum256(float v, float *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// Mean squared error of the normalised outputs and dL/dY = 2 (y - t)/(6 Bt).
__global__ __launch_bounds__(256) void loss_kernel(const float *Yo, const float *Yt,
                                                   int Bt, float gscale, uint16_t *D4,
                                                   uint16_t *D4T, float *loss_part,
                                                   int nblk) {
  __shared__ float red[4];
  const int b = blockIdx.x * 256 + threadIdx.x, mem = blockIdx.y;
  float acc = 0.f;
  if (b < Bt) {
    const float *yo = Yo + (int64_t)mem * Bt * kOutP + (int64_t)b * kOutP;
    const float *yt = Yt + (int64_t)mem * Bt * 8 + (int64_t)b * 8;
    uint16_t *d = D4 + (int64_t)mem * Bt * kInP + (int64_t)b * kInP;
    uint16_t *dt = D4T + (int64_t)mem * kOutP * Bt + b;
    for (int i = 0; i < kOut; ++i) {
      const float r = yo[i] - yt[i];
      acc += r * r;
      const uint16_t u = bf(r * gscale);
      d[i] = u;
      dt[(int64_t)i * Bt] = u;
    }
  }
  const float s = block_sum256(acc, red);
  if (threadIdx.x == 0) loss_part[mem * nblk + blockIdx.x] = s;
}

// Sums loss partials in block order (deterministic) into acc[mem] (+=) and
// out[mem] (=).
__global__ void loss_sum_kernel(const float *loss_part, int nblk, float inv,
                                float *out, float *acc) {
  const int mem = threadIdx.x;
  if (mem >= (int)blockDim.x) return;
  float s = 0.f;
  for (int i = 0; i < nblk; ++i) s += loss_part[mem * nblk + i];
  s *= inv;
  out[mem] = s;
  acc[mem] += s;
}

struct AdamArgs {
  float *P, *Mo, *Ve;
  const float *part[4];
  int64_t sPart[4], sSplit[4];
  int ldp[4];
  int splits;
  uint16_t *w1, *w2, *w3, *w4, *w2t, *w3t, *w4t;
  float *bias;
  float lr, b1, b2, eps, bc1, bc2;
  float *grad;
  int mode;  // 0: Adam update, 1: gradient only, 2: refresh BF16 copies only
};

// One thread per (member, parameter): gradient = sum of the K-split
// partials of its (layer, out, in) entry, in split order.
__global__ void adam_kernel(AdamArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int mem = blockIdx.y;
  if (p >= kNumParams) return;
  int L, o, i;
  bool isb = false;
  if (p < kOffB1) { L = 0; o = p / kIn; i = p % kIn; }
  else if (p < kOffW2) { L = 0; o = p - kOffB1; i = kIn; isb = true; }
  else if (p < kOffB2) { L = 1; o = (p - kOffW2) / kHid; i = (p - kOffW2) % kHid; }
  else if (p < kOffW3) { L = 1; o = p - kOffB2; i = kHid; isb = true; }
  else if (p < kOffB3) { L = 2; o = (p - kOffW3) / kHid; i = (p - kOffW3) % kHid; }
  else if (p < kOffW4) { L = 2; o = p - kOffB3; i = kHid; isb = true; }
  else if (p < kOffB4) { L = 3; o = (p - kOffW4) / kHid; i = (p - kOffW4) % kHid; }
  else { L = 3; o = p - kOffB4; i = kHid; isb = true; }
  const int64_t pi = (int64_t)mem * kNumParams + p;
  float w = a.P[pi];
  if (a.mode != 2) {
    const float *part = a.part[L] + mem * a.sPart[L] + (int64_t)o * a.ldp[L] + i;
    float g = 0.f;
    for (int s = 0; s < a.splits; ++s) g += part[s * a.sSplit[L]];
    if (a.mode == 1) {
      a.grad[pi] = g;
      return;
    }
    const float m = a.b1 * a.Mo[pi] + (1.f - a.b1) * g;
    const float v = a.b2 * a.Ve[pi] + (1.f - a.b2) * g * g;
    a.Mo[pi] = m;
    a.Ve[pi] = v;
    w -= a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
    a.P[pi] = w;
  }
  const uint16_t u = bf(w);
  if (isb) {
    a.bias[mem * kBias + L * kHid + o] = w;
    return;
  }
  switch (L) {
    case 0: a.w1[mem * kW1 + o * kInP + i] = u; break;
    case 1:
      a.w2[mem * kW2 + o * kHid + i] = u;
      a.w2t[mem * kW2 + i * kHid + o] = u;
      break;
    case 2:
      a.w3[mem * kW2 + o * kHid + i] = u;
      a.w3t[mem * kW2 + i * kHid + o] = u;
      break;
    default:
      a.w4[mem * kW4 + o * kHid + i] = u;
      a.w4t[mem * kW4T + i * kInP + o] = u;
      break;
  }
}

// Constant rows of the transposed activation buffers: row `one_row` of
// every member's [rows x Bt] block is 1.
__global__ void ones_row_kernel(uint16_t *T, int64_t member_stride, int one_row, int Bt) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < Bt) T[blockIdx.y * member_stride + (int64_t)one_row * Bt + b] = bf(1.0f);
}

// ---------------------------------------------------------------------------
// Fused inference.  A tile holds TS = 16 Q samples; with JAC it also holds
// six tangent planes (d/d e_j of every activation), so the LDS tile has
// P = Q * planes blocks of 16 rows, one row per (plane, sample), 256 BF16
// features per row.  Each layer computes the transposed product
// W (features) x act^T (rows): wave w owns features [64w, 64w + 64) and lane
// l ends up with 4 consecutive features of one row, written back as one
// 8-byte LDS store.  Plane 0 takes bias and silu, plane j the product with
// silu'(z) of the same (sample, feature), which sits in the same lane.
struct InferArgs {
  const float *X;
  int64_t n;
  int m0, mcount;
  const uint16_t *w1, *w2, *w3, *w4;
  const float *bias;
  const MlpNorm *nm;
  float *Y, *J, *spread;
};

template <int P, int KS>
__device__ __forceinline__ void infer_layer(const uint16_t *act, const uint16_t *W, int ldw,
                                            int w, int lr, int lk, f32x4 (&acc)[P][4]) {
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[p][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < KS; ++kk) {
    bf16x8 wa[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) wa[c] = ld8(W + (64 * w + 16 * c + lr) * ldw + 32 * kk + lk);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const bf16x8 x = ld8(act + (16 * p + lr) * kActLd + 32 * kk + lk);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[p][c] = mfma(wa[c], x, acc[p][c]);
    }
  }
}

template <int Q, int NPL>
__device__ __forceinline__ void infer_hidden_epilogue(uint16_t *act, f32x4 (&acc)[Q * NPL][4],
                                                      const float *bias, int w, int lr,
                                                      int lq) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int f0 = 64 * w + 16 * c + lq;
    const float4 b = *reinterpret_cast<const float4 *>(bias + f0);
#pragma unroll
    for (int sb = 0; sb < Q; ++sb) {
      float h[4], d[4];
      h[0] = silu(acc[sb][c][0] + b.x, d[0]);
      h[1] = silu(acc[sb][c][1] + b.y, d[1]);
      h[2] = silu(acc[sb][c][2] + b.z, d[2]);
      h[3] = silu(acc[sb][c][3] + b.w, d[3]);
      *reinterpret_cast<uint2 *>(act + (16 * sb + lr) * kActLd + f0) =
          pack4(h[0], h[1], h[2], h[3]);
#pragma unroll
      for (int j = 1; j < NPL; ++j) {
        const int p = j * Q + sb;
        *reinterpret_cast<uint2 *>(act + (16 * p + lr) * kActLd + f0) =
            pack4(acc[p][c][0] * d[0], acc[p][c][1] * d[1], acc[p][c][2] * d[2],
                  acc[p][c][3] * d[3]);
      }
    }
  }
}

template <int Q, bool JAC>
__global__ __launch_bounds__(256) void infer_kernel(InferArgs a) {
  constexpr int NPL = JAC ? 7 : 1;
  constexpr int P = Q * NPL;
  constexpr int TS = 16 * Q;
  constexpr int OWN = (P + 3) / 4;
  extern __shared__ uint16_t act[];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = (lane >> 4) * 8, lq = (lane >> 4) * 4;
  const int64_t s0 = (int64_t)blockIdx.x * TS;
  const MlpNorm *nm = a.nm;
  f32x4 fsum[OWN], fsq[OWN];
#pragma unroll
  for (int t = 0; t < OWN; ++t) fsum[t] = fsq[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 acc[P][4];
  for (int mi = 0; mi < a.mcount; ++mi) {
    const int mem = a.m0 + mi;
    // inputs: plane 0 = normalised x (10 of 32 columns), plane j = e_j seed
    for (int e = tid; e < P * 16 * kInP; e += 256) {
      const int row = e / kInP, c = e % kInP;
      const int plane = row / TS, sr = row % TS;
      float v = 0.f;
      if (plane == 0) {
        const int64_t s = s0 + sr;
        if (c < kIn && s < a.n) v = (a.X[s * kIn + c] - nm->in_mean[c]) * nm->in_inv[c];
      } else if (c == plane - 1) {
        v = 1.f;
      }
      act[row * kActLd + c] = bf(v);
    }
    __syncthreads();
    const float *bias = a.bias + mem * kBias;
    infer_layer<P, 1>(act, a.w1 + mem * kW1, kInP, w, lr, lk, acc);
    __syncthreads();
    infer_hidden_epilogue<Q, NPL>(act, acc, bias, w, lr, lq);
    __syncthreads();
    infer_layer<P, 8>(act, a.w2 + mem * kW2, kHid, w, lr, lk, acc);
    __syncthreads();
    infer_hidden_epilogue<Q, NPL>(act, acc, bias + kHid, w, lr, lq);
    __syncthreads();
    infer_layer<P, 8>(act, a.w3 + mem * kW2, kHid, w, lr, lk, acc);
    __syncthreads();
    infer_hidden_epilogue<Q, NPL>(act, acc, bias + 2 * kHid, w, lr, lq);
    __syncthreads();
    // output layer: 16 (6 used) features x 16 rows per block, blocks
    // w, w + 4, ... owned by wave w
    const uint16_t *W4 = a.w4 + mem * kW4;
    const float4 b4 = *reinterpret_cast<const float4 *>(bias + 3 * kHid + lq);
#pragma unroll
    for (int t = 0; t < OWN; ++t) {
      const int p = w + 4 * t;
      if (p < P) {
        f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
          o = mfma(ld8(W4 + lr * kHid + 32 * kk + lk),
                   ld8(act + (16 * p + lr) * kActLd + 32 * kk + lk), o);
        if (p < Q) {
          o[0] += b4.x;
          o[1] += b4.y;
          o[2] += b4.z;
          o[3] += b4.w;
          fsq[t] += o * o;
        }
        fsum[t] += o;
      }
    }
    __syncthreads();
  }
  const float inv_m = 1.0f / (float)a.mcount;
#pragma unroll
  for (int t = 0; t < OWN; ++t) {
    const int p = w + 4 * t;
    if (p >= P) continue;
    const int plane = p / Q, sb = p % Q;
    const int64_t s = s0 + 16 * sb + lr;
    if (s >= a.n) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = lq + q;
      if (i >= kOut) continue;
      const float mean = fsum[t][q] * inv_m;
      if (plane == 0) {
        a.Y[s * kOut + i] = mean * nm->out_std[i] + nm->out_mean[i];
        if (a.spread) {
          const float var = fmaxf(fsq[t][q] * inv_m - mean * mean, 0.f);
          a.spread[s * kOut + i] = sqrtf(var) * nm->out_std[i];
        }
      } else {
        a.J[s * 36 + 6 * i + (plane - 1)] = mean * nm->out_std[i] * nm->in_inv[plane - 1];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Surrogate path helpers.
__device__ inline const double *pick_F(const double *F, int nF, int r) {
  return F + 9 * (int64_t)(nF == 1 ? 0 : r);
}

__global__ void surrogate_inputs_kernel(const double *F, int nF, const float *desc, int n,
                                        float *X) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double *f = pick_F(F, nF, r);
  float *x = X + (int64_t)r * kIn;
  x[0] = (float)(f[0] - 1.0);
  x[1] = (float)(f[4] - 1.0);
  x[2] = (float)(f[8] - 1.0);
  x[3] = (float)(f[5] + f[7]);
  x[4] = (float)(f[2] + f[6]);
  x[5] = (float)(f[1] + f[3]);
  for (int k = 0; k < 4; ++k) x[6 + k] = desc[4 * (int64_t)r + k];
}

__global__ void surrogate_flags_kernel(const float *X, const float *Y, const float *spread,
                                       int n, mf_estimator est, int32_t *flag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int f = 0;
  for (int i = 0; i < kIn; ++i) {
    const float x = X[(int64_t)r * kIn + i];
    const float span = est.in_hi[i] - est.in_lo[i];
    const float pad = isfinite(span) ? est.margin * span : 0.f;  // open bounds stay open
    const float lo = est.in_lo[i] - pad, hi = est.in_hi[i] + pad;
    if (!(x >= lo && x <= hi)) f = 2;  // also catches NaN
  }
  if (f == 0) {
    float smax = 0.f, ymax = 0.f;
    for (int i = 0; i < kOut; ++i) {
      const float y = Y[(int64_t)r * kOut + i];
      ymax = fmaxf(ymax, fabsf(y));
      if (spread) smax = fmaxf(smax, spread[(int64_t)r * kOut + i]);
      if (!isfinite(y)) f = 1;
    }
    if (est.kappa * smax > est.tol * fmaxf(ymax, est.sigma_floor)) f = 1;
  }
  flag[r] = f;
}

// C[:, i] = J e(sym(E_i F)) + (W S - S W), W = skew(E_i F): the derivative
// of the stress along the perturbation F_i(h) F (dF_i/dh at 0 = E_i), with
// the rotation of the stress by the spin of the perturbation.
__global__ void surrogate_outputs_kernel(const double *F, int nF, const float *Y,
                                         const float *J, int n, double *sigma, double *C) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s[6];
  for (int i = 0; i < 6; ++i) s[i] = Y[(int64_t)r * kOut + i];
  for (int i = 0; i < 6; ++i) sigma[(int64_t)r * 6 + i] = s[i];
  if (!C || !J) return;
  const double *f = pick_F(F, nF, r);
  const double S[3][3] = {{s[0], s[5], s[4]}, {s[5], s[1], s[3]}, {s[4], s[3], s[2]}};
  const int vi[6] = {0, 1, 2, 1, 0, 0}, vj[6] = {0, 1, 2, 2, 2, 1};
  for (int col = 0; col < 6; ++col) {
    double E[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (col < 3) {
      E[col][col] = 1.0;
    } else {
      E[vi[col]][vj[col]] = 0.5;
      E[vj[col]][vi[col]] = 0.5;
    }
    double G[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        G[a][b] = E[a][0] * f[b] + E[a][1] * f[3 + b] + E[a][2] * f[6 + b];
    const double de[6] = {G[0][0], G[1][1], G[2][2], G[1][2] + G[2][1], G[0][2] + G[2][0],
                          G[0][1] + G[1][0]};
    double Wm[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Wm[a][b] = 0.5 * (G[a][b] - G[b][a]);
    double R[3][3];  // W S - S W
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double v = 0;
        for (int k = 0; k < 3; ++k) v += Wm[a][k] * S[k][b] - S[a][k] * Wm[k][b];
        R[a][b] = v;
      }
    for (int row = 0; row < 6; ++row) {
      double v = 0;
      for (int k = 0; k < 6; ++k) v += (double)J[(int64_t)r * 36 + 6 * row + k] * de[k];
      v += R[vi[row]][vj[row]];
      C[(int64_t)r * 36 + 6 * row + col] = v;
    }
  }
}

}  // namespace mlpk
}  // namespace mf

using namespace mf;
using namespace mf::mlpk;

struct mf_mlp {
  int K = 1;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf<float> P, Mo, Ve, bias;
  DevBuf<uint16_t> w1, w2, w3, w4, w2t, w3t, w4t;
  DevBuf<MlpNorm> nm_dev;
  MlpNorm nm{};
  float in_std[kIn], out_std_user[kOut];
  int64_t adam_t = 0;
  // training workspace for mini-batches of Bt samples
  int Bt = 0;
  DevBuf<uint16_t> Xn, XnT, H[3], HT[3], S[3], D[3], DT[3], D4, D4T;
  DevBuf<float> Yt, Yo, part[4], loss_part, loss_out, loss_acc, grad;
  DevBuf<int32_t> idx;
  mf_estimator est{};
  double last_ms = 0;
  ~mf_mlp() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

constexpr int kSplitK = 256;  // batch rows per K split of the weight gradients
const int kLdp[4] = {kInP, kTRows, kTRows, kTRows};
const int kPartM[4] = {kHid, kHid, kHid, kOutP};

int upload_norm(mf_mlp *m) {
  MF_HIP(hipMemcpyAsync(m->nm_dev.p, &m->nm, sizeof(MlpNorm), hipMemcpyHostToDevice, m->s));
  return MF_OK;
}

int run_adam(mf_mlp *m, int mode, float lr, float b1, float b2, float eps) {
  AdamArgs a;
  std::memset(&a, 0, sizeof(a));
  a.P = m->P.p;
  a.Mo = m->Mo.p;
  a.Ve = m->Ve.p;
  const int splits = m->Bt > 0 ? m->Bt / kSplitK : 0;
  for (int L = 0; L < 4; ++L) {
    a.part[L] = m->part[L].p;
    a.sSplit[L] = (int64_t)kPartM[L] * kLdp[L];
    a.sPart[L] = a.sSplit[L] * splits;
    a.ldp[L] = kLdp[L];
  }
  a.splits = splits;
  a.w1 = m->w1.p;
  a.w2 = m->w2.p;
  a.w3 = m->w3.p;
  a.w4 = m->w4.p;
  a.w2t = m->w2t.p;
  a.w3t = m->w3t.p;
  a.w4t = m->w4t.p;
  a.bias = m->bias.p;
  a.lr = lr;
  a.b1 = b1;
  a.b2 = b2;
  a.eps = eps;
  if (mode == 0) {
    a.bc1 = (float)(1.0 - std::pow((double)b1, (double)m->adam_t));
    a.bc2 = (float)(1.0 - std::pow((double)b2, (double)m->adam_t));
  }
  a.grad = m->grad.p;
  a.mode = mode;
  hipLaunchKernelGGL(adam_kernel, dim3((kNumParams + 255) / 256, m->K), dim3(256), 0, m->s,
                     a);
  MF_HIP(hipGetLastError());
  return MF_OK;
}

int ensure_workspace(mf_mlp *m, int Bt) {
  if (Bt == m->Bt) return MF_OK;
  m->Bt = 0;
  const size_t K = m->K, B = Bt;
  MF_HIP(m->Xn.alloc(K * B * kInP));
  MF_HIP(m->XnT.alloc(K * kInP * B));
  for (int l = 0; l < 3; ++l) {
    MF_HIP(m->H[l].alloc(K * B * kHid));
    MF_HIP(m->HT[l].alloc(K * kTRows * B));
    MF_HIP(m->S[l].alloc(K * B * kHid));
    MF_HIP(m->D[l].alloc(K * B * kHid));
    MF_HIP(m->DT[l].alloc(K * kHid * B));
  }
  MF_HIP(m->D4.alloc(K * B * kInP));
  MF_HIP(m->D4T.alloc(K * kOutP * B));
  MF_HIP(m->Yt.alloc(K * B * 8));
  MF_HIP(m->Yo.alloc(K * B * kOutP));
  const size_t splits = B / kSplitK;
  for (int L = 0; L < 4; ++L) MF_HIP(m->part[L].alloc(K * splits * kPartM[L] * kLdp[L]));
  MF_HIP(m->loss_part.alloc(K * ((B + 255) / 256)));
  // zero padding (columns / rows that are never written) and ones rows
  MF_HIP(hipMemsetAsync(m->Xn.p, 0, sizeof(uint16_t) * m->Xn.n, m->s));
  MF_HIP(hipMemsetAsync(m->XnT.p, 0, sizeof(uint16_t) * m->XnT.n, m->s));
  MF_HIP(hipMemsetAsync(m->D4.p, 0, sizeof(uint16_t) * m->D4.n, m->s));
  MF_HIP(hipMemsetAsync(m->D4T.p, 0, sizeof(uint16_t) * m->D4T.n, m->s));
  for (int l = 0; l < 3; ++l) {
    MF_HIP(hipMemsetAsync(m->HT[l].p, 0, sizeof(uint16_t) * m->HT[l].n, m->s));
    hipLaunchKernelGGL(ones_row_kernel, dim3((Bt + 255) / 256, m->K), dim3(256), 0, m->s,
                       m->HT[l].p, (int64_t)kTRows * Bt, kHid, Bt);
    MF_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(ones_row_kernel, dim3((Bt + 255) / 256, m->K), dim3(256), 0, m->s,
                     m->XnT.p, (int64_t)kInP * Bt, kIn, Bt);
  MF_HIP(hipGetLastError());
  m->Bt = Bt;
  return MF_OK;
}

template <int EPI>
int gemm(mf_mlp *m, const GemmArgs &g, int splits) {
  const int tiles = ((g.M + 31) / 32) * ((g.N + 31) / 32);
  hipLaunchKernelGGL(gemm_nt<EPI>, dim3((tiles + 3) / 4, splits, m->K), dim3(256), 0, m->s, g);
  MF_HIP(hipGetLastError());
  return MF_OK;
}

GemmArgs gargs(const uint16_t *A, int64_t lda, int64_t sA, const uint16_t *B, int64_t ldb,
               int64_t sB, int M, int N, int kc) {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  g.A = A;
  g.lda = lda;
  g.sA = sA;
  g.B = B;
  g.ldb = ldb;
  g.sB = sB;
  g.M = M;
  g.N = N;
  g.kc = kc;
  return g;
}

// Forward, loss and backward of one mini-batch for every member; leaves the
// weight-gradient partials in m->part and the loss partials in m->loss_part.
int forward_backward(mf_mlp *m, const float *X, const float *Y, int64_t n_data,
                     const int32_t *idx, int64_t idx_stride, int Bt) {
  int e = ensure_workspace(m, Bt);
  if (e != MF_OK) return e;
  const int K = m->K;
  const int64_t B = Bt;
  hipStream_t s = m->s;
  hipLaunchKernelGGL(prep_kernel, dim3((Bt + 255) / 256, K), dim3(256), 0, s, X, Y, n_data,
                     idx, idx_stride, Bt, m->nm_dev.p, m->Xn.p, m->XnT.p, m->Yt.p);
  MF_HIP(hipGetLastError());
  // forward
  const uint16_t *Ain = m->Xn.p;
  int64_t lda = kInP, sA = B * kInP;
  const uint16_t *Wl[3] = {m->w1.p, m->w2.p, m->w3.p};
  const int64_t ldw[3] = {kInP, kHid, kHid}, sW[3] = {kW1, kW2, kW2};
  for (int l = 0; l < 3; ++l) {
    GemmArgs g = gargs(Ain, lda, sA, Wl[l], ldw[l], sW[l], Bt, kHid, (int)ldw[l]);
    g.bias = m->bias.p + l * kHid;
    g.sbias = kBias;
    g.O = m->H[l].p;
    g.OS = m->S[l].p;
    g.OT = m->HT[l].p;
    g.ldo = kHid;
    g.sO = B * kHid;
    g.ldt = B;
    g.sOT = kTRows * B;
    if ((e = gemm<kFwdHidden>(m, g, 1)) != MF_OK) return e;
    Ain = m->H[l].p;
    lda = kHid;
    sA = B * kHid;
  }
  {
    GemmArgs g = gargs(m->H[2].p, kHid, B * kHid, m->w4.p, kHid, kW4, Bt, kOutP, kHid);
    g.bias = m->bias.p + 3 * kHid;
    g.sbias = kBias;
    g.Y = m->Yo.p;
    g.ldy = kOutP;
    g.sY = B * kOutP;
    if ((e = gemm<kFwdOut>(m, g, 1)) != MF_OK) return e;
  }
  const int nblk = (Bt + 255) / 256;
  hipLaunchKernelGGL(loss_kernel, dim3(nblk, K), dim3(256), 0, s, m->Yo.p, m->Yt.p, Bt,
                     2.0f / (6.0f * (float)Bt), m->D4.p, m->D4T.p, m->loss_part.p, nblk);
  MF_HIP(hipGetLastError());
  // backward through the hidden layers: D_l = (D_{l+1} W_{l+1}) * silu'(Z_l)
  {
    const uint16_t *Ad = m->D4.p;
    int64_t ldd = kInP, sD = B * kInP;
    const uint16_t *Wt[3] = {m->w4t.p, m->w3t.p, m->w2t.p};
    const int64_t ldwt[3] = {kInP, kHid, kHid}, sWt[3] = {kW4T, kW2, kW2};
    for (int t = 0; t < 3; ++t) {
      const int l = 2 - t;  // hidden layer receiving the gradient
      GemmArgs g = gargs(Ad, ldd, sD, Wt[t], ldwt[t], sWt[t], Bt, kHid, (int)ldwt[t]);
      g.Sin = m->S[l].p;
      g.O = m->D[l].p;
      g.OT = m->DT[l].p;
      g.ldo = kHid;
      g.sO = B * kHid;
      g.ldt = B;
      g.sOT = kHid * B;
      if ((e = gemm<kBwdHidden>(m, g, 1)) != MF_OK) return e;
      Ad = m->D[l].p;
      ldd = kHid;
      sD = B * kHid;
    }
  }
  // weight gradients (+ bias column): part_L = D_L^T [H_{L-1}^T; 1]
  const int splits = Bt / kSplitK;
  const uint16_t *Ag[4] = {m->DT[0].p, m->DT[1].p, m->DT[2].p, m->D4T.p};
  const int64_t sAg[4] = {kHid * B, kHid * B, kHid * B, kOutP * B};
  const uint16_t *Bg[4] = {m->XnT.p, m->HT[0].p, m->HT[1].p, m->HT[2].p};
  const int64_t sBg[4] = {kInP * B, kTRows * B, kTRows * B, kTRows * B};
  for (int L = 0; L < 4; ++L) {
    GemmArgs g = gargs(Ag[L], B, sAg[L], Bg[L], B, sBg[L], kPartM[L], kLdp[L], kSplitK);
    g.Y = m->part[L].p;
    g.ldy = kLdp[L];
    g.sYsplit = (int64_t)kPartM[L] * kLdp[L];
    g.sY = g.sYsplit * splits;
    if ((e = gemm<kPartial>(m, g, splits)) != MF_OK) return e;
  }
  return MF_OK;
}

bool valid_batch(int64_t Bt) { return Bt > 0 && Bt % MF_MLP_BATCH_ALIGN == 0 && Bt <= (1 << 24); }

}  // namespace

namespace mf {

int mlp_members(const mf_mlp *m) { return m->K; }
const mf_estimator &mlp_estimator(const mf_mlp *m) { return m->est; }

int mlp_predict(mf_mlp *m, int member, const float *X, int64_t n, float *Y, float *J,
                float *spread, hipStream_t s) {
  if (n <= 0) return MF_OK;
  InferArgs a;
  a.X = X;
  a.n = n;
  a.m0 = member < 0 ? 0 : member;
  a.mcount = member < 0 ? m->K : 1;
  a.w1 = m->w1.p;
  a.w2 = m->w2.p;
  a.w3 = m->w3.p;
  a.w4 = m->w4.p;
  a.bias = m->bias.p;
  a.nm = m->nm_dev.p;
  a.Y = Y;
  a.J = J;
  a.spread = spread;
  if (J) {
    constexpr int Q = 1, P = 7;
    const size_t lds = (size_t)P * 16 * kActLd * sizeof(uint16_t);
    hipLaunchKernelGGL((infer_kernel<Q, true>), dim3((unsigned)((n + 16 * Q - 1) / (16 * Q))),
                       dim3(256), lds, s, a);
  } else {
    constexpr int Q = 4, P = 4;
    const size_t lds = (size_t)P * 16 * kActLd * sizeof(uint16_t);
    hipLaunchKernelGGL((infer_kernel<Q, false>), dim3((unsigned)((n + 16 * Q - 1) / (16 * Q))),
                       dim3(256), lds, s, a);
  }
  MF_HIP(hipGetLastError());
  return MF_OK;
}

hipError_t launch_surrogate_inputs(const double *F, int nF, const float *desc, int n, float *X,
                                   hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(surrogate_inputs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, F, nF,
                     desc, n, X);
  return hipGetLastError();
}

hipError_t launch_surrogate_flags(const float *X, const float *Y, const float *spread, int n,
                                  const mf_estimator &est, int32_t *flag, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(surrogate_flags_kernel, dim3((n + 255) / 256), dim3(256), 0, s, X, Y,
                     spread, n, est, flag);
  return hipGetLastError();
}

hipError_t launch_surrogate_outputs(const double *F, int nF, const float *Y, const float *J,
                                    int n, double *sigma, double *C, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(surrogate_outputs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, F, nF,
                     Y, J, n, sigma, C);
  return hipGetLastError();
}

}  // namespace mf

extern "C" {

int32_t mf_mlp_num_params(void) { return kNumParams; }
int32_t mf_mlp_num_members(const mf_mlp *m) { return m ? m->K : 0; }

void mf_default_train_params(mf_train_params *tp) {
  if (!tp) return;
  tp->epochs = 20;
  tp->batch = 4096;
  tp->seed = 1;
  tp->lr = 2e-3f;
  tp->lr_final = 2e-5f;
  tp->beta1 = 0.9f;
  tp->beta2 = 0.999f;
  tp->eps = 1e-8f;
}

int mf_mlp_create(int32_t n_members, uint64_t seed, mf_mlp **out) {
  if (!out || n_members < 1 || n_members > MF_MLP_MAX_MEMBERS) {
    set_error("mf_mlp_create: n_members must be 1..MF_MLP_MAX_MEMBERS");
    return MF_ERR_INVALID;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("mf_mlp_create: no HIP device");
    return MF_ERR_HIP;
  }
  mf_mlp *m = new mf_mlp;
  m->K = n_members;
  auto fail = [&](int e) {
    delete m;
    return e;
  };
  if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess) {
    set_error("mf_mlp_create: stream/event creation failed");
    return fail(MF_ERR_HIP);
  }
  static bool attr_done = false;
  if (!attr_done) {
    const int big = 4 * 16 * kActLd * (int)sizeof(uint16_t);
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&infer_kernel<4, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, big) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void *>(&infer_kernel<1, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            7 * 16 * kActLd * (int)sizeof(uint16_t)) != hipSuccess) {
      set_error("mf_mlp_create: LDS attribute");
      return fail(MF_ERR_HIP);
    }
    attr_done = true;
  }
  const size_t K = n_members;
  std::vector<float> P(K * kNumParams);
  // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
  for (size_t k = 0; k < K; ++k) {
    SplitMix64 rng(seed, 1000 + k);
    float *w = P.data() + k * kNumParams;
    auto fill = [&](int off, int cnt, int fan_in) {
      const double b = 1.0 / std::sqrt((double)fan_in);
      for (int i = 0; i < cnt; ++i) w[off + i] = (float)((2.0 * rng.uniform() - 1.0) * b);
    };
    fill(kOffW1, kHid * kIn, kIn);
    fill(kOffB1, kHid, kIn);
    fill(kOffW2, kHid * kHid, kHid);
    fill(kOffB2, kHid, kHid);
    fill(kOffW3, kHid * kHid, kHid);
    fill(kOffB3, kHid, kHid);
    fill(kOffW4, kOut * kHid, kHid);
    fill(kOffB4, kOut, kHid);
  }
  int e = MF_OK;
  auto chk = [&](hipError_t h) {
    if (h != hipSuccess && e == MF_OK) {
      set_error(std::string("mf_mlp_create: ") + hipGetErrorString(h));
      e = MF_ERR_HIP;
    }
  };
  chk(m->P.upload(P, m->s));
  chk(m->Mo.alloc(P.size()));
  chk(m->Ve.alloc(P.size()));
  chk(m->bias.alloc(K * kBias));
  chk(m->w1.alloc(K * kW1));
  chk(m->w2.alloc(K * kW2));
  chk(m->w3.alloc(K * kW2));
  chk(m->w4.alloc(K * kW4));
  chk(m->w2t.alloc(K * kW2));
  chk(m->w3t.alloc(K * kW2));
  chk(m->w4t.alloc(K * kW4T));
  chk(m->nm_dev.alloc(1));
  chk(m->grad.alloc(K * kNumParams));
  chk(m->loss_out.alloc(K));
  chk(m->loss_acc.alloc(K));
  if (e != MF_OK) return fail(e);
  chk(hipMemsetAsync(m->Mo.p, 0, sizeof(float) * m->Mo.n, m->s));
  chk(hipMemsetAsync(m->Ve.p, 0, sizeof(float) * m->Ve.n, m->s));
  chk(hipMemsetAsync(m->bias.p, 0, sizeof(float) * m->bias.n, m->s));
  chk(hipMemsetAsync(m->w1.p, 0, sizeof(uint16_t) * m->w1.n, m->s));
  chk(hipMemsetAsync(m->w4.p, 0, sizeof(uint16_t) * m->w4.n, m->s));
  chk(hipMemsetAsync(m->w4t.p, 0, sizeof(uint16_t) * m->w4t.n, m->s));
  if (e != MF_OK) return fail(e);
  for (int i = 0; i < kIn; ++i) {
    m->nm.in_mean[i] = 0.f;
    m->nm.in_inv[i] = 1.f;
    m->in_std[i] = 1.f;
    m->est.in_lo[i] = -INFINITY;
    m->est.in_hi[i] = INFINITY;
  }
  for (int i = 0; i < 8; ++i) {
    m->nm.out_mean[i] = 0.f;
    m->nm.out_std[i] = 1.f;
    m->nm.out_inv[i] = 1.f;
  }
  for (int i = 0; i < kOut; ++i) m->out_std_user[i] = 1.f;
  m->est.margin = 0.f;
  m->est.kappa = 1.f;
  m->est.sigma_floor = 0.f;
  m->est.tol = INFINITY;
  if ((e = upload_norm(m)) != MF_OK) return fail(e);
  if ((e = run_adam(m, 2, 0, 0, 0, 0)) != MF_OK) return fail(e);
  if (hipStreamSynchronize(m->s) != hipSuccess) {
    set_error("mf_mlp_create: initialisation failed");
    return fail(MF_ERR_HIP);
  }
  *out = m;
  return MF_OK;
}

void mf_mlp_destroy(mf_mlp *m) { delete m; }

int mf_mlp_set_normalization(mf_mlp *m, const float *in_mean, const float *in_std,
                             const float *out_mean, const float *out_std) {
  if (!m || !in_mean || !in_std || !out_mean || !out_std) {
    set_error("mf_mlp_set_normalization: null argument");
    return MF_ERR_INVALID;
  }
  for (int i = 0; i < kIn; ++i) {
    const float sd = in_std[i] > 0.f ? in_std[i] : 1.f;
    m->nm.in_mean[i] = in_mean[i];
    m->in_std[i] = sd;
    m->nm.in_inv[i] = (float)(1.0 / (double)sd);
  }
  for (int i = 0; i < kOut; ++i) {
    const float sd = out_std[i] > 0.f ? out_std[i] : 1.f;
    m->nm.out_mean[i] = out_mean[i];
    m->nm.out_std[i] = sd;
    m->out_std_user[i] = sd;
    m->nm.out_inv[i] = (float)(1.0 / (double)sd);
  }
  int e = upload_norm(m);
  if (e != MF_OK) return e;
  MF_HIP(hipStreamSynchronize(m->s));
  return MF_OK;
}

int mf_mlp_get_normalization(const mf_mlp *m, float *in_mean, float *in_std, float *out_mean,
                             float *out_std) {
  if (!m) return MF_ERR_INVALID;
  for (int i = 0; i < kIn; ++i) {
    if (in_mean) in_mean[i] = m->nm.in_mean[i];
    if (in_std) in_std[i] = m->in_std[i];
  }
  for (int i = 0; i < kOut; ++i) {
    if (out_mean) out_mean[i] = m->nm.out_mean[i];
    if (out_std) out_std[i] = m->nm.out_std[i];
  }
  return MF_OK;
}

int mf_mlp_get_weights(mf_mlp *m, int32_t member, float *w) {
  if (!m || !w || member < 0 || member >= m->K) {
    set_error("mf_mlp_get_weights: bad member or buffer");
    return MF_ERR_INVALID;
  }
  MF_HIP(hipMemcpyAsync(w, m->P.p + (size_t)member * kNumParams, sizeof(float) * kNumParams,
                        hipMemcpyDeviceToHost, m->s));
  MF_HIP(hipStreamSynchronize(m->s));
  return MF_OK;
}

int mf_mlp_set_weights(mf_mlp *m, int32_t member, const float *w) {
  if (!m || !w || member < 0 || member >= m->K) {
    set_error("mf_mlp_set_weights: bad member or buffer");
    return MF_ERR_INVALID;
  }
  const size_t off = (size_t)member * kNumParams;
  MF_HIP(hipMemcpyAsync(m->P.p + off, w, sizeof(float) * kNumParams, hipMemcpyHostToDevice,
                        m->s));
  MF_HIP(hipMemsetAsync(m->Mo.p + off, 0, sizeof(float) * kNumParams, m->s));
  MF_HIP(hipMemsetAsync(m->Ve.p + off, 0, sizeof(float) * kNumParams, m->s));
  int e = run_adam(m, 2, 0, 0, 0, 0);
  if (e != MF_OK) return e;
  MF_HIP(hipStreamSynchronize(m->s));
  return MF_OK;
}

int mf_mlp_train_step_device(mf_mlp *m, const float *X_dev, const float *Y_dev,
                             int64_t n_data, const int32_t *idx_dev, int32_t batch, float lr,
                             float *loss) {
  if (!m || !X_dev || !Y_dev || n_data <= 0 || !valid_batch(batch) ||
      (!idx_dev && batch > n_data)) {
    set_error("mf_mlp_train_step_device: bad arguments (batch must be a positive multiple of "
              "MF_MLP_BATCH_ALIGN)");
    return MF_ERR_INVALID;
  }
  int e = forward_backward(m, X_dev, Y_dev, n_data, idx_dev, batch, batch);
  if (e != MF_OK) return e;
  m->adam_t += 1;
  if ((e = run_adam(m, 0, lr, 0.9f, 0.999f, 1e-8f)) != MF_OK) return e;
  if (loss) {
    const int nblk = (batch + 255) / 256;
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(m->K), 0, m->s, m->loss_part.p, nblk,
                       1.0f / (6.0f * (float)batch), m->loss_out.p, m->loss_acc.p);
    MF_HIP(hipGetLastError());
    std::vector<float> l(m->K);
    MF_HIP(hipMemcpyAsync(l.data(), m->loss_out.p, sizeof(float) * m->K,
                          hipMemcpyDeviceToHost, m->s));
    MF_HIP(hipStreamSynchronize(m->s));
    double sum = 0;
    for (float v : l) sum += v;
    *loss = (float)(sum / m->K);
  }
  return MF_OK;
}

int mf_mlp_fit_device(mf_mlp *m, const float *X_dev, const float *Y_dev, int64_t n,
                      const mf_train_params *tp, float *epoch_loss) {
  if (!m || !X_dev || !Y_dev || !tp || tp->epochs < 0 || !valid_batch(tp->batch) ||
      n < tp->batch || n > (int64_t)1 << 30) {
    set_error("mf_mlp_fit_device: bad arguments");
    return MF_ERR_INVALID;
  }
  const int K = m->K, Bt = tp->batch;
  const int64_t steps = n / Bt;
  const int64_t total = steps * tp->epochs;
  std::vector<int32_t> perm((size_t)K * n);
  MF_HIP(m->idx.ensure(perm.size()));
  const int nblk = (Bt + 255) / 256;
  int64_t it = 0;
  for (int ep = 0; ep < tp->epochs; ++ep) {
    MF_HIP(hipStreamSynchronize(m->s));  // previous epoch done with idx
    for (int k = 0; k < K; ++k) {
      int32_t *p = perm.data() + (size_t)k * n;
      for (int64_t i = 0; i < n; ++i) p[i] = (int32_t)i;
      SplitMix64 rng(tp->seed, (uint64_t)ep * K + k);
      for (int64_t i = n - 1; i > 0; --i) {
        const int64_t j = (int64_t)(rng.next() % (uint64_t)(i + 1));
        std::swap(p[i], p[j]);
      }
    }
    MF_HIP(hipMemcpyAsync(m->idx.p, perm.data(), sizeof(int32_t) * perm.size(),
                          hipMemcpyHostToDevice, m->s));
    MF_HIP(hipMemsetAsync(m->loss_acc.p, 0, sizeof(float) * K, m->s));
    for (int64_t st = 0; st < steps; ++st, ++it) {
      const double frac = total > 1 ? (double)it / (double)(total - 1) : 1.0;
      const float lr = (float)(tp->lr_final + 0.5 * (tp->lr - tp->lr_final) *
                                                  (1.0 + std::cos(3.14159265358979323846 * frac)));
      int e = forward_backward(m, X_dev, Y_dev, n, m->idx.p + st * Bt, n, Bt);
      if (e != MF_OK) return e;
      m->adam_t += 1;
      if ((e = run_adam(m, 0, lr, tp->beta1, tp->beta2, tp->eps)) != MF_OK) return e;
      hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(K), 0, m->s, m->loss_part.p, nblk,
                         1.0f / (6.0f * (float)Bt), m->loss_out.p, m->loss_acc.p);
      MF_HIP(hipGetLastError());
    }
    if (epoch_loss) {
      std::vector<float> l(K);
      MF_HIP(hipMemcpyAsync(l.data(), m->loss_acc.p, sizeof(float) * K, hipMemcpyDeviceToHost,
                            m->s));
      MF_HIP(hipStreamSynchronize(m->s));
      for (int k = 0; k < K; ++k)
        epoch_loss[(size_t)ep * K + k] = steps > 0 ? l[k] / (float)steps : 0.f;
    }
  }
  MF_HIP(hipStreamSynchronize(m->s));
  return MF_OK;
}

int mf_mlp_gradient_device(mf_mlp *m, int32_t member, const float *X_dev, const float *Y_dev,
                           int32_t n, float *grad_host, float *loss) {
  if (!m || member < 0 || member >= m->K || !X_dev || !Y_dev || !grad_host ||
      !valid_batch(n)) {
    set_error("mf_mlp_gradient_device: bad arguments");
    return MF_ERR_INVALID;
  }
  int e = forward_backward(m, X_dev, Y_dev, n, nullptr, 0, n);
  if (e != MF_OK) return e;
  if ((e = run_adam(m, 1, 0, 0, 0, 0)) != MF_OK) return e;
  const int nblk = (n + 255) / 256;
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(m->K), 0, m->s, m->loss_part.p, nblk,
                     1.0f / (6.0f * (float)n), m->loss_out.p, m->loss_acc.p);
  MF_HIP(hipGetLastError());
  MF_HIP(hipMemcpyAsync(grad_host, m->grad.p + (size_t)member * kNumParams,
                        sizeof(float) * kNumParams, hipMemcpyDeviceToHost, m->s));
  float l = 0;
  MF_HIP(hipMemcpyAsync(&l, m->loss_out.p + member, sizeof(float), hipMemcpyDeviceToHost,
                        m->s));
  MF_HIP(hipStreamSynchronize(m->s));
  if (loss) *loss = l;
  return MF_OK;
}

int mf_mlp_predict_device(mf_mlp *m, int32_t member, const float *X_dev, int64_t n,
                          float *Y_dev, float *J_dev, float *spread_dev) {
  if (!m || member < -1 || member >= m->K || n < 0 || (n > 0 && (!X_dev || !Y_dev)) ||
      n > ((int64_t)1 << 31) * 16) {
    set_error("mf_mlp_predict_device: bad arguments");
    return MF_ERR_INVALID;
  }
  MF_HIP(hipEventRecord(m->ev0, m->s));
  int e = mlp_predict(m, member, X_dev, n, Y_dev, J_dev, spread_dev, m->s);
  if (e != MF_OK) return e;
  MF_HIP(hipEventRecord(m->ev1, m->s));
  MF_HIP(hipEventSynchronize(m->ev1));
  float ms = 0;
  MF_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  m->last_ms = ms;
  return MF_OK;
}

int mf_mlp_last_predict_ms(const mf_mlp *m, double *ms) {
  if (!m || !ms) return MF_ERR_INVALID;
  *ms = m->last_ms;
  return MF_OK;
}

void *mf_mlp_stream(mf_mlp *m) { return m ? (void *)m->s : nullptr; }

int mf_mlp_set_estimator(mf_mlp *m, const mf_estimator *est) {
  if (!m || !est || !(est->margin >= 0.f) || !(est->kappa >= 0.f) ||
      !(est->sigma_floor >= 0.f) || !(est->tol >= 0.f)) {
    set_error("mf_mlp_set_estimator: bad arguments");
    return MF_ERR_INVALID;
  }
  m->est = *est;
  return MF_OK;
}

int mf_mlp_get_estimator(const mf_mlp *m, mf_estimator *est) {
  if (!m || !est) return MF_ERR_INVALID;
  *est = m->est;
  return MF_OK;
}

}  // extern "C"
