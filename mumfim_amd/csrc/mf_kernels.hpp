// Host-visible interface of the HIP kernels: argument blocks, kernel
// variants and their LDS footprints.
#pragma once
#include <hip/hip_runtime.h>

#include "mf_internal.hpp"

namespace mf {

// Per-solve flags.
enum : int32_t {
  kWriteX = 1,  // write final current coordinates to x_out
  kWriteU = 2,  // write u = x - X to u_out
};

struct SolveArgs {
  const RveDev *rves;
  const double *X;          // 3 per node
  const double *mass;       // per node
  const int32_t *deg;       // per node
  const int32_t *slice_base, *slice_D;
  const int32_t *ell_w;
  const double *ell_l0;     // FP64 kernels
  const float *ell_il;      // FP32 kernels: 1/l0
  const int32_t *fib;       // 2 per fiber (kernel ranks)
  const int32_t *fstream;   // 4 per fiber: ranks a, b and their ELL entries
  const int32_t *ell_fref;  // per ELL entry: fiber record index | sign bit
  const float *fparam;      // 4 per fiber: reference vector, law coefficient
  const ReactDev *reacts;
  const double *x_in, *u_in;
  double *x_out, *u_out;
  const double *V;          // per RVE
  const int32_t *solve_rve;
  const double *solve_F;    // 9 per solve
  const int32_t *solve_flags;
  const int32_t *order;     // blockIdx -> solve
  double *sigma;            // 6 per solve
  int32_t *status;          // per solve
  int32_t *steps;           // per solve (diagnostic)
  int32_t mode;
  int32_t n_steps, max_steps, freq;
  double eta, crit, tol;
  // per-RVE viscous damping (BatchedFiberRVEAnalysisExplicit.h:213-218);
  // NULL: eta for every RVE
  const double *eta_rve;
  int32_t amplitude;        // MF_AMP_* (ramp mode)
  // one-RVE-per-wave variants: bytes of LDS per wave (each wave's region
  // starts at wave * wave_lds)
  int32_t wave_lds;
  double energy_eps;        // > 0: energy balance check (energy variants)
  double *energy;           // 4 per solve (energy variants), may be NULL
  // global-state variants: per-solve scratch at gstate + block * gstride
  double *gstate;
  int64_t gstride;          // doubles per solve
  // Step chunks (ramp mode, register kernel).  The launch runs
  // n_chunks * group_n workgroups: workgroup i runs chunk i / group_n of the
  // group's solve i % group_n, steps [c * chunk_steps, (c + 1) * chunk_steps)
  // (the last chunk to n_steps).  Between chunks a solve's lane state lives at
  // cstate + (i % group_n) * cstride floats, and cflag[solve] counts its
  // finished chunks (release/acquire at agent scope; the last chunk resets it
  // to 0, so every launch starts from zeros).  n_chunks <= 1: no chunks.
  int32_t n_chunks, chunk_steps, group_n, pad3_;
  float *cstate;
  int64_t cstride;          // floats per solve
  int32_t *cflag;
  // tangent-linear variants: direction of each solve's perturbation of F
  // (9 per solve) and the stress derivative along it (6 per solve)
  const double *solve_G;
  double *dsigma;
};

// Floats of chunk state per solve of a register-kernel workgroup of B
// threads and NPT node slots per thread: (u hi, u lo, v) x 3 per slot and
// lane, then the step size and the residual.
constexpr int64_t chunk_stride(int B, int NPT) { return ((int64_t)NPT * 9 * B + 4 + 63) & ~63; }
// Spin limit of a chunk waiting for its predecessor (s_sleep 127 per try,
// ~3.4 us): ~14 s, far beyond any chunk; a solve that runs out is reported NaN.
constexpr int kChunkSpinLimit = 1 << 22;

// A kernel variant: threads per solve, free-node slots per thread and the
// placement flags.
enum : int32_t {
  kEllInLds = 1,    // free-node incidence lists staged in LDS
  kDoubleBuf = 2,   // positions double-buffered in LDS (one barrier per step)
  kEllInRegs = 4,   // incidence lists in registers, one free node per thread
  kLinearOnly = 8,  // every fiber of the RVE uses one linear reaction
  kEnergy = 16,     // FP64 with energy diagnostics
  kMix2 = 32,       // at most one linear and one nonlinear reaction (FP32c registers)
  kFiberPar = 64,   // FP32c fiber-parallel: lanes own fibers, nodes sum end forces
  kFiberStream = 128,  // FP32c fiber-streamed: per-fiber forces, nodes gather by reference
  kWaveRve = 4096,     // FP32c registers, one RVE per wave, block / 64 RVEs per workgroup
  kGlobalState = 8192, // FP64 state in global memory: any RVE size
  kFiberOnce = 16384,  // FP64, t = f/l once per fiber, nodes gather by reference
  kTangentLinear = 32768,  // FP32c solve with its derivative along solve_G
};
// Stride (float4 entries) of the two position buffers of the register
// kernel: buffer 1 starts 32 KiB into LDS, a compile-time offset the LDS
// loads take as an immediate; RVEs of up to this many nodes.
constexpr int kRegBufStride = 2048;
// The same stride of the one-RVE-per-wave variants, whose LDS region per
// wave holds RVEs of up to this many nodes.
constexpr int kWaveBufStride = 384;
// Fibers per thread of a fiber-parallel variant (flags bits 8-11; the
// other flags stay clear of them).
constexpr int kFplShift = 8;
inline int variant_fpl(int32_t flags) { return (flags >> kFplShift) & 15; }



struct Variant {
  int32_t prec;    // MF_PREC_*
  int32_t block;   // threads per workgroup
  int32_t npt;     // free-node slots per thread
  int32_t flags;
};

// All compiled variants (the planner picks from this list).
const Variant *variants(int *count);
// LDS bytes a variant needs for one RVE (N nodes, Nb boundary, M fibers).
size_t lds_bytes(const Variant &v, int N, int Nb, int M, int ell_free_entries);
// Largest dynamic LDS a workgroup may request.
constexpr size_t kMaxLds = 160 * 1024;

hipError_t launch_solve(const Variant &v, const SolveArgs &a, int n_blocks,
                        size_t lds, hipStream_t s);
// Workgroups of variant v (with `lds` dynamic LDS bytes) resident per CU;
// 0 when v has no step-chunk support.
int chunk_occupancy(const Variant &v, size_t lds);
hipError_t lds_probe(double *bytes_per_s);
hipError_t hbm_probe(int64_t bytes, double *bytes_per_s);
hipError_t launch_broadcast_F(const double *F, double *out, int n, hipStream_t s);
hipError_t launch_update_volume(const double *F, int nF, double *V, int n_rve,
                                hipStream_t s);
hipError_t launch_fd(const double *sig_a, const double *sig_b, double *C,
                     int n_rve, int column, double inv_h, int stride_a,
                     int stride_b, hipStream_t s);
hipError_t launch_fd_all(const double *sigma, double *C, int n_rve, int scheme,
                         double h, hipStream_t s);
// C of n_rve RVEs from tangent-linear solves 6 r + i (d sigma along i).
hipError_t launch_tl_assemble(const double *dsigma, double *C, int n_rve, hipStream_t s);
hipError_t launch_orientation(const RveDev *rves, const double *x,
                              const int32_t *fib, const double *normal,
                              int n_normal, double *omega, int n_rve,
                              hipStream_t s);

}  // namespace mf
