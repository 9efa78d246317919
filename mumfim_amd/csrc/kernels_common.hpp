// Batched explicit fiber-network RVE solve for gfx950.
//
// One workgroup runs one solve (one RVE under one F) for all of its steps.
// Node positions live in LDS (double-buffered where they fit, so a step needs
// a single barrier); each thread owns fixed free-node slots whose position,
// velocity and force stay in registers for the whole solve.  A free node
// gathers the forces of its incident fibers from a sliced-ELL incidence list
// (64 degree-sorted nodes per slice = one wavefront, entries in increasing
// fiber index), which replaces the reference's atomic scatter
// (BatchedFiberRVEAnalysisExplicitBase.h:226-252) and reproduces its
// summation order exactly.
//
// solve64: FP64 with the reference's operation order and no FMA contraction
//          (BatchedFiberRVEAnalysisExplicitTeamOuterLoop.h:180-437).
// solve32c: FP32 stretch/force/update arithmetic on the algebraic form
//          t = f/l = EA/(2 l0) (lambda^2 - 1), positions carried as
//          compensated (hi, lo) FP32 pairs updated with TwoSum, stress
//          reduced in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include "mf_kernels.hpp"

namespace mf {
namespace kern {

constexpr int kWave = 64;
constexpr int kRedStride = 8;  // doubles per wave in the reduction scratch

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
// a . b with the contraction spelt out (x product rounded once, y and z
// fused): every kernel shape then rounds a fiber's stretch, a squared length
// or a residual the same way, whatever the compiler's own contraction choice
// in the surrounding code
__device__ inline float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
  return fmaf(az, bz, fmaf(ay, by, ax * bx));
}

__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, kWave);
    v = w < v ? w : v;
  }
  return v;
}

// Sums NV values over the workgroup; every thread receives the totals.  The
// caller guarantees a barrier separates this call from the previous read of
// `red`.
template <int B, int NV>
__device__ inline void block_sum(double (&v)[NV], double *red) {
  constexpr int NW = B / kWave;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) red[w * kRedStride + i] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) s += red[k * kRedStride + i];
    v[i] = s;
  }
}

// Ramp-mode boundary amplitude after step k (1-based) of S:
// linear xi = k/S, or SmoothAmp xi^3 (10 - 15 xi + 6 xi^2)
// (ExplicitAmplitude.h:26-40), evaluated in the reference's operation order.
__device__ inline double ramp_alpha(int k, int S, int amplitude) {
  const double xi = (double)k / (double)S;
  if (amplitude == MF_AMP_SMOOTHSTEP) return xi * xi * xi * (10 - 15 * xi + 6 * xi * xi);
  return xi;
}


// Boundary record of the FP32 kernels (48 bytes of LDS, 16-byte aligned):
// the start displacement u0 and the ramp increment g = (F - I) x_b in FP64
// (u0x u0y u0z gx gy gz), read as three 16-byte pairs.  The step's position
// u0 + alpha g is formed in FP64 and rounded once to FP32: a boundary
// displacement is ~0.03 while a boundary fiber stretches by ~5e-3, so an ulp
// of the position is ~4e-7 of the stretch, and an FP32 evaluation (three
// roundings) measured 3-4x the median stress error of the rounded FP64 one.
__device__ inline void bnd_set(double *rec, const double (&u0)[3], const double (&g)[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rec[k] = u0[k];
    rec[3 + k] = g[k];
  }
}

// u_b = u0 + alpha g in FP64.
__device__ inline void bnd_disp(const double *rec, double alpha, double (&ub)[3]) {
  const double2 *p = reinterpret_cast<const double2 *>(rec);
  const double2 q0 = p[0], q1 = p[1], q2 = p[2];
  ub[0] = fma(alpha, q1.y, q0.x);
  ub[1] = fma(alpha, q2.x, q0.y);
  ub[2] = fma(alpha, q2.y, q1.x);
}

// The step's boundary displacement, rounded once to FP32.
__device__ inline void bnd_step(const double *rec, double alpha, float (&ub)[3]) {
  double u[3];
  bnd_disp(rec, alpha, u);
#pragma unroll
  for (int k = 0; k < 3; ++k) ub[k] = (float)u[k];
}

// Ramp factor of the FP32 kernels' step loop: xi = k * (1/S) is within an
// ulp of k / S and saves an FP64 division per step (the positions it moves
// are rounded to FP32 anyway); the final stress uses ramp_alpha.
__device__ inline double ramp_alpha_step(int k, double inv_S, int amplitude) {
  const double xi = (double)k * inv_S;
  if (amplitude == MF_AMP_SMOOTHSTEP) return xi * xi * xi * (10 - 15 * xi + 6 * xi * xi);
  return xi;
}

// min of a (slot 0) and sum of b (slot 1) in one pass.
template <int B>
__device__ inline void block_min_sum(double &a, double &b, double *red) {
  constexpr int NW = B / kWave;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  double m = wave_min(a), s = wave_sum(b);
  if (lane == 0) {
    red[w * kRedStride + 0] = m;
    red[w * kRedStride + 1] = s;
  }
  __syncthreads();
  m = red[0];
  s = red[1];
#pragma unroll
  for (int k = 1; k < NW; ++k) {
    double mk = red[k * kRedStride + 0];
    m = mk < m ? mk : m;
    s += red[k * kRedStride + 1];
  }
  a = m;
  b = s;
}

__device__ inline size_t align16(size_t b) { return (b + 15) & ~size_t(15); }

// whole 16-byte loads of records in LDS through generic pointers
typedef float ld16_f4 __attribute__((ext_vector_type(4), aligned(16)));
typedef int ld16_i4 __attribute__((ext_vector_type(4), aligned(16)));
__device__ inline float4 ld16(const float4 *p) {
  const ld16_f4 v = *reinterpret_cast<const ld16_f4 *>(p);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline int4 ld16(const int4 *p) {
  const ld16_i4 v = *reinterpret_cast<const ld16_i4 *>(p);
  return make_int4(v.x, v.y, v.z, v.w);
}

// ---- reaction laws, FP64 reference formulas (FiberReactions.cc:5-71) ------
__device__ inline double nl_force(double alpha, double B, double lr,
                                  double g) {
  return alpha * lr * (exp(B * g) - 1);
}
__device__ inline double nl_deriv(double alpha, double B, double lr, double l0,
                                  double g) {
  return alpha / l0 * ((1 + lr * lr * B) * exp(B * g) - 1);
}
__device__ inline double react64(const ReactDev &q, double l0, double l) {
#pragma clang fp contract(off)
  double lr = l / l0;
  if (q.type == MF_REACTION_LINEAR) {
    double g = 1.0 / 2.0 * (lr * lr - 1);
    return lr * q.E * q.A * g;
  }
  double alpha = q.E * q.A / q.B;
  if (lr > q.lt) {
    double gt = 1.0 / 2.0 * (q.lt * q.lt - 1);
    double m = nl_deriv(alpha, q.B, q.lt, l0, gt);
    double b = nl_force(alpha, q.B, q.lt, gt);
    return m * (lr - q.lt) * l0 + b;
  } else if (lr < q.lc) {
    double gc = 1.0 / 2.0 * (q.lc * q.lc - 1);
    double m = nl_deriv(alpha, q.B, q.lc, l0, gc);
    double b = nl_force(alpha, q.B, q.lc, gc);
    return m * (lr - q.lc) * l0 + b;
  }
  double g = 1.0 / 2.0 * (lr * lr - 1);
  return nl_force(alpha, q.B, lr, g);
}

// ---- reaction laws, FP32 algebraic forms --------------------------------
// t = f/l as a function of e2 = lambda^2 - 1 = (l^2 - l0^2)/l0^2, which the
// FP32 kernel forms without cancellation from the displacement difference.
// Linear: t = EA/2 il e2.  Toe: t = alpha il expm1(B/2 e2).  Beyond a
// transition: f = K (lam - L0) + b, so t = il (K - C / lam) with
// C = K L0 - b folded on the host and 1/lam = rsqrt(1 + e2): one rsqrt and
// two operations, as accurate in FP32 as the direct form (both lose the
// same absolute eps K near the transition).
__device__ inline float ext_law(float K, float C, float il, float ilam) {
  return il * fmaf(-C, ilam, K);
}

// Branch-free toe/extension law over a reaction table repacked in LDS as
// three float4 per reaction:
//   [0] EA/2, alpha = EA/B, B/2, 1 if nonlinear else 0
//   [1] lt^2 - 1, Kt, Ct, 0         (tension extension, lambda > lt)
//   [2] lc^2 - 1, Kc, Cc, 0         (compression extension, lambda < lc)
// In the toe region x = B/2 (lambda^2 - 1) lies in [B/2 (lc^2-1), B/2 (lt^2-1)];
// clamping e2 to that range keeps the polynomial's argument inside it, and
// the selected branch ignores the clamped value outside the region.
constexpr float kToePolyMax = 0.25f;  // |x| bound of the degree-7 expm1
constexpr int kReactTabStride = 3;    // float4 per reaction

// Law of a run of list entries in the kernels that stream incidence lists:
// the one linear reaction, the nonlinear-first entries of a
// one-linear-one-nonlinear RVE (either law per lane), its later entries
// (linear on every lane), or the packed table.
enum { kLawLin = 0, kLawMixNl = 1, kLawMixLin = 2, kLawTab = 3 };
// Fourth word of a raw buffer resource on gfx9-family GPUs: 32-bit data
// format (a zero format makes untyped buffer loads return 0).
constexpr int kBufRsrcWord3 = 0x00020000;

__device__ inline void pack_react_table(const ReactDev &q, float4 *t) {
  const bool nl = q.type == MF_REACTION_NONLINEAR;
  t[0] = make_float4(q.EA_half, nl ? q.alpha : 0.0f, nl ? q.B_half : 0.0f, nl ? 1.0f : 0.0f);
  t[1] = make_float4(nl ? q.lt2m1 : 0.0f, q.Kt, q.Ct, 0.0f);
  t[2] = make_float4(nl ? q.lc2m1 : 0.0f, q.Kc, q.Cc, 0.0f);
}

// expm1(x) for |x| <= kToePolyMax: Taylor series to x^7 (truncation below
// x^8/8! = 3.8e-10 relative at the bound).
__device__ inline float expm1_poly(float x) {
  float p = 1.0f / 5040.0f;
  p = p * x + 1.0f / 720.0f;
  p = p * x + 1.0f / 120.0f;
  p = p * x + 1.0f / 24.0f;
  p = p * x + 1.0f / 6.0f;
  p = p * x + 0.5f;
  p = p * x + 1.0f;
  return p * x;
}

// alpha expm1(B/2 y) for |B/2 y| <= kToePolyMax, as y P(y) with the
// host-folded coefficients alpha (B/2)^k / k! (expm1_poly's Taylor series
// with the argument and result scalings taken into the coefficients)
__device__ inline float toe_poly(const ReactDev &q, float y) {
  float p = q.pe[6];
  p = p * y + q.pe[5];
  p = p * y + q.pe[4];
  p = p * y + q.pe[3];
  p = p * y + q.pe[2];
  p = p * y + q.pe[1];
  p = p * y + q.pe[0];
  return p * y;
}

// Branch-free: every lane evaluates both laws and both toe forms and selects.
// expm1 on |x| <= kToePolyMax is the polynomial; beyond it exp(x) - 1 loses
// at most a few ulp (e^x - 1 > 0.28 there).
__device__ inline float react32_tab(const float4 *T, unsigned rid, float e2, float il) {
  const float4 a = T[kReactTabStride * rid];
  const float4 b = T[kReactTabStride * rid + 1], c = T[kReactTabStride * rid + 2];
  const float tl = a.x * il * e2;
  const bool hi = e2 > b.x, lo = e2 < c.x;
  const float x = a.z * fminf(fmaxf(e2, c.x), b.x);
  const float em = fabsf(x) <= kToePolyMax ? expm1_poly(x) : __expf(x) - 1.0f;
  const float tt = a.y * il * em;
  const float te = ext_law(hi ? b.y : c.y, hi ? b.z : c.z, il, __frsqrt_rn(1.0f + e2));
  const float tn = (hi || lo) ? te : tt;
  return a.w != 0.0f ? tn : tl;
}

// t of a nonlinear reaction whose toe range lies within the polynomial's
// bound, branch-free: the toe polynomial and the extension are both
// evaluated and one is selected.  Outside the toe range the polynomial's
// value is dropped, so its argument needs no clamp.
__device__ inline float toe_ext_t(const ReactDev &nl, float e2, float il) {
  const bool hi = e2 > nl.lt2m1, lo = e2 < nl.lc2m1;
  const float tt = il * toe_poly(nl, e2);
  const float te = ext_law(hi ? nl.Kt : nl.Kc, hi ? nl.Ct : nl.Cc, il, __frsqrt_rn(1.0f + e2));
  return (hi || lo) ? te : tt;
}

// One linear (EA/2 = ea_half_lin) and one nonlinear reaction (nl) with a
// toe range within the polynomial's bound: t for a fiber of either law.
__device__ inline float react32_mix2(const ReactDev &nl, float ea_half_lin, bool is_nl,
                                     float e2, float il) {
  return is_nl ? toe_ext_t(nl, e2, il) : ea_half_lin * il * e2;
}

// t and dt/de2 of react32_tab (tangent-linear kernel).  Linear: EA/2 il;
// toe: d/de2 alpha il expm1(B/2 e2) = B/2 (alpha il + t); beyond a
// transition t = il (K - C / lam) with lam = sqrt(1 + e2), whose derivative
// is il C / (2 lam^3).
__device__ inline float react32_tab_d(const float4 *T, unsigned rid, float e2, float il,
                                      float &dt) {
  const float4 a = T[kReactTabStride * rid];
  const float4 b = T[kReactTabStride * rid + 1], c = T[kReactTabStride * rid + 2];
  const float tl = a.x * il * e2;
  const bool hi = e2 > b.x, lo = e2 < c.x;
  const float x = a.z * fminf(fmaxf(e2, c.x), b.x);
  const float em = fabsf(x) <= kToePolyMax ? expm1_poly(x) : __expf(x) - 1.0f;
  const float tt = a.y * il * em;
  const float ilam = __frsqrt_rn(1.0f + e2);
  const float C = hi ? b.z : c.z;
  const float te = ext_law(hi ? b.y : c.y, C, il, ilam);
  const float dte = 0.5f * C * il * (ilam * ilam * ilam);
  const bool ext = hi || lo;
  dt = a.w != 0.0f ? (ext ? dte : a.z * (a.y * il + tt)) : a.x * il;
  return a.w != 0.0f ? (ext ? te : tt) : tl;
}

// t and dt/de2 of react32_mix2.
__device__ inline float react32_mix2_d(const ReactDev &nl, float ea_half_lin, bool is_nl,
                                       float e2, float il, float &dt) {
  const bool hi = e2 > nl.lt2m1, lo = e2 < nl.lc2m1;
  const float tt = il * toe_poly(nl, e2);
  const float ilam = __frsqrt_rn(1.0f + e2);
  const float C = hi ? nl.Ct : nl.Cc;
  const float te = ext_law(hi ? nl.Kt : nl.Kc, C, il, ilam);
  const float dte = 0.5f * C * il * (ilam * ilam * ilam);
  const bool ext = hi || lo;
  dt = is_nl ? (ext ? dte : nl.B_half * (nl.alpha * il + tt)) : ea_half_lin * il;
  return is_nl ? (ext ? te : tt) : ea_half_lin * il * e2;
}

// The reaction table qualifies for react32_mix2 (uniform over the
// workgroup); nl_rid and ea_half_lin describe it.
__device__ inline bool mix2_table(const ReactDev *q, int n, int &nl_rid, float &ea_half_lin) {
  if (n < 1 || n > 2) return false;
  int n_nl = 0;
  nl_rid = -1;
  ea_half_lin = 0.0f;
  for (int i = 0; i < n; ++i) {
    if (q[i].type == MF_REACTION_NONLINEAR) {
      ++n_nl;
      nl_rid = i;
      if (!(q[i].B_half * fmaxf(fabsf(q[i].lt2m1), fabsf(q[i].lc2m1)) <= kToePolyMax))
        return false;
    } else {
      ea_half_lin = q[i].EA_half;
    }
  }
  return n_nl == 1;
}

// Fast path when the whole RVE uses one linear reaction.
struct Lin64 {
  double E, A;
};

}  // namespace kern
}  // namespace mf

// Tangent-linear FP32c variants: (threads, free-node slots per thread).
#define MF_TL_LIST(X) X(256, 1) X(512, 1) X(1024, 1) X(1024, 2) X(1024, 3)

// Fiber-streamed FP32c variants (large RVEs): (threads, fibers per thread,
// free nodes per thread, reaction mode flag kLinearOnly|kMix2).  The widest
// 1024-thread ones spill a few registers in the step loop (7 toe-law or 8
// linear fibers per thread) and still beat the LDS-list kernel; 8 toe-law
// fibers per thread did not.  At 768 threads (168 VGPRs) 7 fibers and 3
// nodes per thread fit without spills and run 1900-2100-node toe-law RVEs
// 2% faster than 8 and 4.
// One-RVE-per-wave FP32c register variants: (RVEs per workgroup, free-node
// slots per thread), each in the linear and one-linear-one-nonlinear modes
// (four slots per thread spill registers).
#define MF_WAVE_LIST(X) X(1, 3) X(4, 3)

#define MF_FIBER_STREAM_LIST(X)                                              \
  X(512, 4, 2, kLinearOnly) X(512, 4, 2, kMix2)                              \
  X(1024, 4, 2, kLinearOnly) X(1024, 4, 2, kMix2)                            \
  X(1024, 6, 3, kLinearOnly) X(1024, 6, 3, kMix2)                            \
  X(1024, 7, 3, kLinearOnly) X(1024, 7, 3, kMix2)                            \
  X(1024, 8, 3, kLinearOnly)                                                 \
  X(768, 6, 3, kLinearOnly) X(768, 6, 3, kMix2)                              \
  X(768, 7, 3, kLinearOnly) X(768, 7, 3, kMix2)                              \
  X(768, 8, 4, kLinearOnly) X(768, 8, 4, kMix2)                              \
  X(768, 10, 4, kLinearOnly) X(768, 10, 4, kMix2)

// Fiber-parallel FP32c variants: (threads, fibers per thread, free nodes per
// thread, reaction mode flags kLinearOnly|kMix2, 0 = reaction table).
#define MF_FIBER_PAR_LIST(X)                                                 \
  X(512, 2, 1, kLinearOnly) X(512, 2, 1, kMix2) X(512, 2, 1, 0)              \
  X(512, 3, 1, kLinearOnly) X(512, 3, 1, kMix2) X(512, 3, 1, 0)              \
  X(768, 2, 1, kLinearOnly) X(768, 2, 1, kMix2)                              \
  X(1024, 2, 1, kLinearOnly) X(1024, 2, 1, kMix2)                            \
  X(1024, 3, 2, kLinearOnly) X(1024, 3, 2, kMix2)                            \
  X(1024, 4, 2, kLinearOnly) X(1024, 4, 2, kMix2)

// Compiled kernel variants: (precision, threads, free-node slots per thread,
// flags kEllInLds|kDoubleBuf).
#define MF_VARIANT_LIST(X)                                                   \
  X(MF_PREC_FP64, 256, 1, 3) X(MF_PREC_FP64, 256, 1, 2)                      \
  X(MF_PREC_FP64, 256, 2, 3) X(MF_PREC_FP64, 256, 2, 2)                      \
  X(MF_PREC_FP64, 256, 4, 3) X(MF_PREC_FP64, 256, 4, 2)                      \
  X(MF_PREC_FP64, 256, 4, 0) X(MF_PREC_FP64, 512, 2, 3)                      \
  X(MF_PREC_FP64, 512, 2, 2) X(MF_PREC_FP64, 512, 4, 2)                      \
  X(MF_PREC_FP64, 512, 4, 0) X(MF_PREC_FP64, 1024, 2, 2)                     \
  X(MF_PREC_FP64, 1024, 2, 0) X(MF_PREC_FP64, 1024, 4, 0)                    \
  X(MF_PREC_FP64, 512, 1, 3) X(MF_PREC_FP64, 512, 1, 2)                      \
  X(MF_PREC_FP64, 1024, 1, 2)                                                \
  X(MF_PREC_FP32C, 256, 1, 3) X(MF_PREC_FP32C, 256, 1, 2)                    \
  X(MF_PREC_FP32C, 256, 2, 3) X(MF_PREC_FP32C, 256, 2, 2)                    \
  X(MF_PREC_FP32C, 256, 4, 3) X(MF_PREC_FP32C, 256, 4, 2)                    \
  X(MF_PREC_FP32C, 256, 4, 0) X(MF_PREC_FP32C, 512, 2, 3)                    \
  X(MF_PREC_FP32C, 512, 2, 2) X(MF_PREC_FP32C, 512, 4, 3)                    \
  X(MF_PREC_FP32C, 512, 4, 2) X(MF_PREC_FP32C, 512, 4, 0)                    \
  X(MF_PREC_FP32C, 1024, 2, 2) X(MF_PREC_FP32C, 1024, 2, 0)                  \
  X(MF_PREC_FP32C, 1024, 4, 2) X(MF_PREC_FP32C, 1024, 4, 0)

