// Small FP64 kernels around the solve (volume update, finite-difference
// tangent columns, orientation tensors) and the kernel-variant table.
// Compiled with -ffp-contract=off.
#include <algorithm>

#include "kernels_common.hpp"

namespace mf {
namespace kern {

// ---- small kernels -------------------------------------------------------------
// LDS read-bandwidth probe.  Per iteration each thread reads 16 float4 at
// one lane-consecutive (conflict-free) base with immediate offsets 4 KiB
// apart over a 64 KiB tile; the base rotates by a wave width every iteration
// so no load is loop-invariant.  Every dword is folded with 3-input integer
// adds (v_add3_u32), two VALU instructions per 1 KiB wave read, half the
// issue slots a 256 B/clk LDS leaves per read, so the probe is bound by the
// LDS, not by issue.
__global__ __launch_bounds__(256) void lds_probe_kernel(uint32_t *out, int iters) {
  __shared__ float4 tile[4096];
  const int t = threadIdx.x;
  for (int i = t; i < 4096; i += 256) tile[i] = make_float4(i, 1.0f, 2.0f, 3.0f);
  __syncthreads();
  uint32_t a0 = 0, a1 = 0;
  for (int it = 0; it < iters; ++it) {
    const float4 *p = tile + ((it * 64 + t) & 255);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float4 v = p[k * 256];
      a0 = a0 + __float_as_uint(v.x) + __float_as_uint(v.y);
      a1 = a1 + __float_as_uint(v.z) + __float_as_uint(v.w);
    }
  }
  out[blockIdx.x * 256 + t] = a0 + a1;
}

__global__ void copy_probe_kernel(const float4 *in, float4 *out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

__global__ void broadcast_F_kernel(const double *F, double *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 9 * n) out[i] = F[i % 9];
}

// V <- det(F) V (updateVolume, ...Base.h:127-154)
__global__ void update_volume_kernel(const double *F, int nF, double *V,
                                     int n) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *dg = F + 9 * (nF == 1 ? 0 : i);
  double J = dg[0] * (dg[4] * dg[8] - dg[5] * dg[7]) -
             dg[1] * (dg[3] * dg[8] - dg[5] * dg[6]) +
             dg[2] * (dg[3] * dg[7] - dg[4] * dg[6]);
  V[i] = J * V[i];
}

// Finite-difference tangent columns in FP64.  sigma holds, per RVE, the
// solves [base, col0(+), col1(+), ..., col5(+)] (forward) or
// [base, col0(+), col0(-), ..., col5(+), col5(-)] (central), 6 values each.
__global__ void fd_kernel(const double *sigma, double *C, int n, int scheme,
                          double h) {
#pragma clang fp contract(off)
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * 36) return;
  const int r = idx / 36, j = (idx % 36) / 6, i = idx % 6;
  const int per = scheme == MF_TANGENT_CENTRAL ? 13 : 7;
  const double *s = sigma + (int64_t)r * per * 6;
  double v;
  if (scheme == MF_TANGENT_CENTRAL)
    v = (s[(1 + 2 * i) * 6 + j] - s[(2 + 2 * i) * 6 + j]) / (2 * h);
  else
    v = (s[(1 + i) * 6 + j] - s[j]) / h;
  C[(int64_t)r * 36 + j * 6 + i] = v;
}

// Tangent-linear columns: solve 6 r + i holds d sigma along direction i.
__global__ void tl_assemble_kernel(const double *dsigma, double *C, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * 36) return;
  const int r = idx / 36, j = (idx % 36) / 6, i = idx % 6;
  C[(int64_t)r * 36 + j * 6 + i] = dsigma[((int64_t)r * 6 + i) * 6 + j];
}

// One FD column with separate base/perturbed arrays (relax-mode sequence,
// BatchedFiberRVEAnalysisExplicit.h:394-402).
__global__ void fd_column_kernel(const double *sig1, const double *sig0,
                                 double *C, int n, int column, double h,
                                 int s1, int s0) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int j = 0; j < 6; ++j)
    C[(int64_t)r * 36 + j * 6 + column] =
        (sig1[(int64_t)r * s1 + j] - sig0[(int64_t)r * s0 + j]) / h;
}

// Orientation tensors (BatchedComputeOrientationTensor.h:56-176), one
// workgroup per RVE, FP64 block reduction.
template <int B>
__global__ __launch_bounds__(B) void orientation_kernel(
    const RveDev *rves, const double *x, const int32_t *fib,
    const double *normal, int n_normal, double *omega) {
#pragma clang fp contract(off)
  __shared__ double red[(B / kWave) * kRedStride * 2];
  const int rv = blockIdx.x;
  const RveDev R = rves[rv];
  double nrm[3] = {0, 0, 0};
  if (normal) {
    const double *q = normal + 3 * (n_normal == 1 ? 0 : rv);
    double n = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    nrm[0] = q[0] / n;
    nrm[1] = q[1] / n;
    nrm[2] = q[2] / n;
  }
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int32_t *fb = fib + 2 * R.fib_off;
  const double *xs = x + 3 * R.node_off;
  for (int e = threadIdx.x; e < R.M; e += B) {
    const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
    double lx = xs[3 * n2] - xs[3 * n1];
    double ly = xs[3 * n2 + 1] - xs[3 * n1 + 1];
    double lz = xs[3 * n2 + 2] - xs[3 * n1 + 2];
    double l = sqrt(lx * lx + ly * ly + lz * lz);
    lx = lx / l;
    ly = ly / l;
    lz = lz / l;
    if (normal) {
      double dotn = lx * nrm[0] + ly * nrm[1] + lz * nrm[2];
      lx = lx - nrm[0] * dotn;
      ly = ly - nrm[1] * dotn;
      lz = lz - nrm[2] * dotn;
      l = sqrt(lx * lx + ly * ly + lz * lz);
      lx = lx / l;
      ly = ly / l;
      lz = lz / l;
    }
    const double v[3] = {lx, ly, lz};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) acc[3 * i + j] += v[i] * v[j];
  }
  double a6[6] = {acc[0], acc[1], acc[2], acc[3], acc[4], acc[5]};
  double b3[3] = {acc[6], acc[7], acc[8]};
  block_sum<B, 6>(a6, red);
  block_sum<B, 3>(b3, red + (B / kWave) * kRedStride);
  if (threadIdx.x == 0) {
    double *o = omega + 9 * rv;
    for (int i = 0; i < 6; ++i) o[i] = a6[i] / R.M;
    for (int i = 0; i < 3; ++i) o[6 + i] = b3[i] / R.M;
  }
}


}  // namespace kern

#define MF_VARIANT_ENTRY(P, B, N, F) {P, B, N, F},
// FP64 with energy diagnostics; the list's FP32c entries become inert
// placeholders (precision -1 matches no plan)
#define MF_ENERGY_ENTRY(P, B, N, F) {P == MF_PREC_FP64 ? P : -1, B, N, F | kEnergy},
static const Variant kVariants[] = {
    MF_VARIANT_LIST(MF_VARIANT_ENTRY)
    MF_VARIANT_LIST(MF_ENERGY_ENTRY)
    // FP32c with incidence lists in registers, one free node per thread:
    // linear-only and one-linear-one-nonlinear modes, reaction-table mode at
    // 256 and 512 threads
#define MF_REG2(B)                                                 \
    {MF_PREC_FP32C, B, 1, kEllInRegs | kDoubleBuf | kLinearOnly},  \
    {MF_PREC_FP32C, B, 1, kEllInRegs | kDoubleBuf | kMix2},
    MF_REG2(256) MF_REG2(512) MF_REG2(640) MF_REG2(768) MF_REG2(1024)
#undef MF_REG2
    // one RVE per wave, W RVEs per workgroup: (W, slots per thread)
#define MF_WAVE_ENTRY(W, N)                                                        \
    {MF_PREC_FP32C, 64 * (W), N, kEllInRegs | kDoubleBuf | kLinearOnly | kWaveRve}, \
    {MF_PREC_FP32C, 64 * (W), N, kEllInRegs | kDoubleBuf | kMix2 | kWaveRve},
    MF_WAVE_LIST(MF_WAVE_ENTRY)
#undef MF_WAVE_ENTRY
    {MF_PREC_FP32C, 256, 1, kEllInRegs | kDoubleBuf},
    {MF_PREC_FP32C, 512, 1, kEllInRegs | kDoubleBuf},
    // several free nodes per thread
    {MF_PREC_FP32C, 768, 2, kEllInRegs | kDoubleBuf | kLinearOnly},
    {MF_PREC_FP32C, 256, 2, kEllInRegs | kDoubleBuf | kLinearOnly},
    {MF_PREC_FP32C, 512, 3, kEllInRegs | kDoubleBuf | kLinearOnly},
    {MF_PREC_FP32C, 512, 3, kEllInRegs | kDoubleBuf | kMix2},
    {MF_PREC_FP32C, 512, 4, kEllInRegs | kDoubleBuf | kLinearOnly},
    // fiber-parallel: (threads, fibers per thread, free nodes per thread)
#define MF_FPAR_ENTRY(B, FPL, NPL, FL) {MF_PREC_FP32C, B, NPL, kFiberPar | FL | ((FPL) << kFplShift)},
    MF_FIBER_PAR_LIST(MF_FPAR_ENTRY)
#undef MF_FPAR_ENTRY
    // fiber-streamed: (threads, fibers per thread, free nodes per thread)
#define MF_FSTR_ENTRY(B, FPL, NPL, FL) \
  {MF_PREC_FP32C, B, NPL, kFiberStream | FL | ((FPL) << kFplShift)},
    MF_FIBER_STREAM_LIST(MF_FSTR_ENTRY)
#undef MF_FSTR_ENTRY
    // FP64, each fiber evaluated once per step: (threads, free nodes per thread)
    {MF_PREC_FP64, 256, 1, kFiberOnce},
    {MF_PREC_FP64, 256, 2, kFiberOnce},
    {MF_PREC_FP64, 512, 1, kFiberOnce},
    {MF_PREC_FP64, 512, 2, kFiberOnce},
    {MF_PREC_FP64, 1024, 1, kFiberOnce},
    {MF_PREC_FP64, 1024, 2, kFiberOnce},
    // FP64 with the state in global memory: RVEs no other variant holds
    {MF_PREC_FP64, 1024, 1, kGlobalState},
    // FP32c tangent-linear: (threads, free nodes per thread)
#define MF_TL_ENTRY(B, N) {MF_PREC_FP32C, B, N, kTangentLinear},
    MF_TL_LIST(MF_TL_ENTRY)
#undef MF_TL_ENTRY
};

const Variant *variants(int *count) {
  *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
}

size_t lds_bytes(const Variant &v, int N, int Nb, int M, int ell_free_entries) {
  const size_t Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const size_t NW = (v.flags & kWaveRve) ? 1 : v.block / kern::kWave;  // waves per RVE
  size_t b = 0;
  if (v.flags & kFiberOnce) {
    // double-buffered positions, t per fiber (M bounds the fibers with a
    // free end), scratch, boundary state, reactions
    return sizeof(double) * 6 * Np + sizeof(double) * (((size_t)M + 1) & ~size_t(1)) +
           sizeof(double) * NW * kern::kRedStride + sizeof(double) * 6 * Nbp +
           ((sizeof(ReactDev) * kMaxReactions + 15) & ~size_t(15));
  }
  if (v.flags & kTangentLinear) {
    // displacements, their derivatives, reference coordinates; scratch;
    // boundary state (FP64) and its derivative (FP32); packed table
    return 48 * Np + sizeof(double) * NW * kern::kRedStride + sizeof(double) * 6 * Nbp +
           16 * Nbp + sizeof(float) * 4 * 3 * kMaxReactions;
  }
  if (v.flags & kGlobalState) {
    // scratch, boundary state, reactions (positions etc. in global memory)
    return sizeof(double) * NW * kern::kRedStride + sizeof(double) * 6 * Nbp +
           ((sizeof(ReactDev) * kMaxReactions + 15) & ~size_t(15));
  }
  if (v.flags & kFiberStream) {
    // 12-byte displacement and force records (+ zero record), scratch,
    // boundary state, packed table
    const size_t nff = (std::max<size_t>((size_t)M + 1, (size_t)N) + 1) & ~size_t(1);
    b = (12 * (Np + nff) + 15) & ~size_t(15);
    b += sizeof(double) * NW * kern::kRedStride + sizeof(double) * 6 * Nbp;
    b = (b + 15) & ~size_t(15);
    return b + sizeof(float) * 4 * 3 * kMaxReactions;
  }
  if (v.prec == MF_PREC_FP64) {
    b += sizeof(double) * 3 * Np * ((v.flags & kDoubleBuf) ? 2 : 1);
  } else if (v.flags & kFiberPar) {
    // displacements + end-force slots (a spare per lane), which hold the
    // low parts at the end
    b += 16 * (Np + std::max<size_t>((size_t)ell_free_entries + kern::kWave, Np));
  } else if (v.flags & kEllInRegs) {
    // displacement buffers only, fixed stride (per RVE: with one RVE per
    // wave the launch multiplies this by the RVEs per workgroup)
    b += 16 * (((v.flags & kWaveRve) ? kWaveBufStride : kRegBufStride) + Np);
  } else {
    // displacement buffers + reference coordinates
    b += 16 * Np * ((v.flags & kDoubleBuf) ? 3 : 2);
  }
  b += sizeof(double) * NW * kern::kRedStride;
  b += sizeof(double) * 6 * Nbp;
  if (v.flags & kEnergy) b += sizeof(double) * 3 * Nbp;
  b += v.prec == MF_PREC_FP64 ? (sizeof(ReactDev) * kMaxReactions + 15) & ~size_t(15)
                              : sizeof(float) * 4 * 3 * kMaxReactions;  // packed table
  if (v.flags & kEllInLds)
    b += (size_t)ell_free_entries *
         (sizeof(int32_t) + (v.prec == MF_PREC_FP64 ? sizeof(double) : sizeof(float)));
  return b;
}

hipError_t launch_solve64(const Variant &v, const SolveArgs &a, int n_blocks,
                          size_t lds, hipStream_t s);
hipError_t launch_solve32(const Variant &v, const SolveArgs &a, int n_blocks,
                          size_t lds, hipStream_t s);
hipError_t launch_solve32t(const Variant &v, const SolveArgs &a, int n_blocks,
                           size_t lds, hipStream_t s);

hipError_t launch_solve(const Variant &v, const SolveArgs &a, int n_blocks,
                        size_t lds, hipStream_t s) {
  if (n_blocks <= 0) return hipSuccess;
  if (v.flags & kTangentLinear) return launch_solve32t(v, a, n_blocks, lds, s);
  return v.prec == MF_PREC_FP64 ? launch_solve64(v, a, n_blocks, lds, s)
                                : launch_solve32(v, a, n_blocks, lds, s);
}

hipError_t lds_probe(double *bytes_per_s) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return hipErrorNoDevice;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const int blocks = cus * 8, iters = 4096;
  uint32_t *out = nullptr;
  e = hipMalloc(&out, sizeof(uint32_t) * blocks * 256);
  if (e != hipSuccess) return e;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  hipLaunchKernelGGL(kern::lds_probe_kernel, dim3(blocks), dim3(256), 0, 0, out, 64);
  (void)hipEventRecord(a, 0);
  hipLaunchKernelGGL(kern::lds_probe_kernel, dim3(blocks), dim3(256), 0, 0, out, iters);
  (void)hipEventRecord(b, 0);
  e = hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  *bytes_per_s = (double)blocks * 256 * iters * 16 * 16 / (ms * 1e-3);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  return e;
}

hipError_t hbm_probe(int64_t bytes, double *bytes_per_s) {
  const int64_t n = bytes / 16;
  float4 *in = nullptr, *out = nullptr;
  hipError_t e = hipMalloc(&in, n * 16);
  if (e != hipSuccess) return e;
  e = hipMalloc(&out, n * 16);
  if (e != hipSuccess) {
    (void)hipFree(in);
    return e;
  }
  (void)hipMemset(in, 0, n * 16);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const int blocks = 4096;
  hipLaunchKernelGGL(kern::copy_probe_kernel, dim3(blocks), dim3(256), 0, 0, in, out, n);
  (void)hipEventRecord(a, 0);
  const int reps = 5;
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(kern::copy_probe_kernel, dim3(blocks), dim3(256), 0, 0, in, out, n);
  (void)hipEventRecord(b, 0);
  e = hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  *bytes_per_s = 2.0 * n * 16 * reps / (ms * 1e-3);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(in);
  (void)hipFree(out);
  return e;
}

hipError_t launch_broadcast_F(const double *F, double *out, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern::broadcast_F_kernel, dim3((9 * n + 255) / 256), dim3(256), 0,
                     s, F, out, n);
  return hipGetLastError();
}

hipError_t launch_update_volume(const double *F, int nF, double *V, int n,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern::update_volume_kernel, dim3((n + 255) / 256), dim3(256), 0,
                     s, F, nF, V, n);
  return hipGetLastError();
}

hipError_t launch_fd_all(const double *sigma, double *C, int n, int scheme,
                         double h, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern::fd_kernel, dim3((n * 36 + 255) / 256), dim3(256), 0, s,
                     sigma, C, n, scheme, h);
  return hipGetLastError();
}

hipError_t launch_tl_assemble(const double *dsigma, double *C, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern::tl_assemble_kernel, dim3((n * 36 + 255) / 256), dim3(256), 0, s,
                     dsigma, C, n);
  return hipGetLastError();
}

hipError_t launch_fd(const double *sig_a, const double *sig_b, double *C,
                     int n, int column, double h, int stride_a, int stride_b,
                     hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern::fd_column_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                     sig_a, sig_b, C, n, column, h, stride_a, stride_b);
  return hipGetLastError();
}

hipError_t launch_orientation(const RveDev *rves, const double *x,
                              const int32_t *fib, const double *normal,
                              int n_normal, double *omega, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern::orientation_kernel<256>, dim3(n), dim3(256), 0, s, rves,
                     x, fib, normal, n_normal, omega);
  return hipGetLastError();
}

}  // namespace mf
