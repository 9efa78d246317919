// C ABI of the batched RVE analysis: construction (packing + upload), the
// run/tangent state machine of BatchedFiberRVEAnalysisExplicit
// (src/mumfim/microscale/BatchedFiberRVEAnalysisExplicit.h:80-416) and the
// launch planner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "mf_internal.hpp"
#include "mf_kernels.hpp"
#include "dev_util.hpp"
#include "mlp_internal.hpp"

namespace mf {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace mf

using namespace mf;

namespace {

void make_react_dev(const Reaction &q, ReactDev &d) {
  std::memset(&d, 0, sizeof(d));
  d.type = q.type;
  d.E = q.E;
  d.A = q.area();
  d.B = q.B;
  d.lt = q.lt;
  d.lc = q.lc;
  const double EA = q.E * d.A;
  d.EA_half = (float)(0.5 * EA);
  if (q.type == MF_REACTION_NONLINEAR) {
    const double alpha = EA / q.B;
    d.alpha = (float)alpha;
    d.B_half = (float)(0.5 * q.B);
    auto ext = [&](double lam, float &K, float &C) {
      // f(lam) beyond a transition: m*(lam - lt)*l0 + b with m*l0 and b
      // independent of l0 (FiberReactions.cc:16-62); t = f/l = il (K - C/lam)
      double g = 0.5 * (lam * lam - 1);
      const double k = alpha * ((1 + lam * lam * q.B) * std::exp(q.B * g) - 1);
      const double b = alpha * lam * (std::exp(q.B * g) - 1);
      K = (float)k;
      C = (float)(k * lam - b);
    };
    ext(q.lt, d.Kt, d.Ct);
    ext(q.lc, d.Kc, d.Cc);
    d.lt2m1 = (float)(q.lt * q.lt - 1.0);
    d.lc2m1 = (float)(q.lc * q.lc - 1.0);
    double c = alpha;
    for (int k = 1; k <= 7; ++k) {
      c *= 0.5 * q.B / k;
      d.pe[k - 1] = (float)c;
    }
  }
}

// Almansi-strain perturbation F_i = sqrt((I - 2e_i)^-1) with e_i the unit
// strain of Voigt component i scaled by h (engineering shear), as in
// BatchedFiberRVEAnalysisExplicit.h:340-389 and RVEAnalysis.cc:15-64.
// Rounded without contraction, operation for operation as the oracle does,
// so the FP64 tangent's perturbed solves see the same F bits.
void perturbation(int i, double h, double Fp[9]) {
#pragma clang fp contract(off)
  double D1 = std::sqrt(1.0 / (1 - 2 * h));
  const double l1 = 1;
  const double l2 = (1 + h) / (1 - h * h);
  const double l3 = (1 - h) / (1 - h * h);
  double l2pl3 = 0.5 * (std::sqrt(l2) + std::sqrt(l3));
  double l2ml3 = 0.5 * (std::sqrt(l2) - std::sqrt(l3));
  for (int k = 0; k < 9; ++k) Fp[k] = 0;
  switch (i) {
    case 0: Fp[0] = D1; Fp[4] = 1; Fp[8] = 1; break;
    case 1: Fp[0] = 1; Fp[4] = D1; Fp[8] = 1; break;
    case 2: Fp[0] = 1; Fp[4] = 1; Fp[8] = D1; break;
    case 3: Fp[0] = l1; Fp[4] = l2pl3; Fp[8] = l2pl3; Fp[5] = l2ml3; Fp[7] = l2ml3; break;
    case 4: Fp[0] = l2pl3; Fp[4] = l1; Fp[8] = l2pl3; Fp[2] = l2ml3; Fp[6] = l2ml3; break;
    default: Fp[0] = l2pl3; Fp[4] = l2pl3; Fp[8] = l1; Fp[1] = l2ml3; Fp[3] = l2ml3; break;
  }
}

void matmul3(const double A[9], const double B[9], double C[9]) {
#pragma clang fp contract(off)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * k + j];
      C[3 * i + j] = s;
    }
}

// d P_i / d h at h = 0 of the perturbation above: the unit Voigt strain of
// direction i (engineering shear: 1/2 off the diagonal).
void perturbation_rate(int i, double E[9]) {
  for (int k = 0; k < 9; ++k) E[k] = 0;
  switch (i) {
    case 0: E[0] = 1; break;
    case 1: E[4] = 1; break;
    case 2: E[8] = 1; break;
    case 3: E[5] = E[7] = 0.5; break;
    case 4: E[2] = E[6] = 0.5; break;
    default: E[1] = E[3] = 0.5; break;
  }
}

bool params_valid(const mf_solver_params &p) {
  return (p.mode == MF_MODE_RELAX || p.mode == MF_MODE_RAMP) &&
         (p.precision == MF_PREC_FP64 || p.precision == MF_PREC_FP32C) &&
         (p.tangent_precision == MF_PREC_FP64 ||
          p.tangent_precision == MF_PREC_FP32C) &&
         (p.reaction_policy == MF_REACT_BATCHED_REFERENCE ||
          p.reaction_policy == MF_REACT_PER_FIBER) &&
         (p.tangent_scheme == MF_TANGENT_FORWARD || p.tangent_scheme == MF_TANGENT_CENTRAL ||
          (p.tangent_scheme == MF_TANGENT_LINEAR && p.mode == MF_MODE_RAMP &&
           p.tangent_precision == MF_PREC_FP32C)) &&
         p.dt_update_freq >= 1 && p.max_steps >= 1 &&
         (p.mode != MF_MODE_RAMP || p.n_steps >= 1) && p.fd_step > 0 &&
         p.fd_step < 0.25 && p.crit_time_scale_factor > 0 &&
         p.visc_damp_coeff >= 0 &&
         (p.amplitude == MF_AMP_LINEAR || p.amplitude == MF_AMP_SMOOTHSTEP) &&
         (p.energy_diagnostics == 0 ||
          (p.energy_diagnostics == 1 && p.precision == MF_PREC_FP64)) &&
         p.energy_check_eps >= 0 &&
         (p.fp32_kernel == MF_FP32_KERNEL_AUTO || p.fp32_kernel == MF_FP32_KERNEL_NODE ||
          p.fp32_kernel == MF_FP32_KERNEL_FIBER || p.fp32_kernel == MF_FP32_KERNEL_WAVE) &&
         (p.block_size == 0 || p.block_size == 64 || p.block_size == 256 || p.block_size == 512 ||
          p.block_size == 640 || p.block_size == 768 || p.block_size == 1024) &&
         p.chunks >= 0 && p.chunks <= 16;
}

template <class F>
void parallel_for(int n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4 || nt == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> ts;
  std::atomic<int> next{0};
  for (unsigned t = 0; t < nt; ++t)
    ts.emplace_back([&] {
      for (int i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &t : ts) t.join();
}

}  // namespace

// Autotuner bucket (see tune_key).
struct TuneKey {
  int prec, lin, regok, lanes64, n256;
  bool operator<(const TuneKey &o) const {
    return std::tie(prec, lin, regok, lanes64, n256) <
           std::tie(o.prec, o.lin, o.regok, o.lanes64, o.n256);
  }
  bool operator==(const TuneKey &o) const {
    return std::tie(prec, lin, regok, lanes64, n256) ==
           std::tie(o.prec, o.lin, o.regok, o.lanes64, o.n256);
  }
};

struct mf_batch {
  mf_solver_params p{};
  int R = 0;
  int policy = 0;
  std::vector<PackedRVE> host;
  std::vector<RveDev> rdev;
  int64_t total_nodes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // extra streams so launch groups of different kernel variants overlap
  static constexpr int kAux = 3;
  hipStream_t aux[kAux] = {nullptr, nullptr, nullptr};
  hipEvent_t aux_done[kAux] = {nullptr, nullptr, nullptr};
  float last_ms = 0;
  int last_launches = 0;
  DevBuf<RveDev> rves;
  DevBuf<double> X, mass, l0, V, cur_stress, F_in;
  DevBuf<int32_t> deg, slice_base, slice_D, ell_w, fib, fstream, ell_fref;
  DevBuf<double> ell_l0;
  DevBuf<float> ell_il, fparam;
  DevBuf<ReactDev> reacts;
  DevBuf<double> xs[2], us[2];
  int cur = 0;
  bool have_run = false;   // current_stress_ valid
  bool have_prev = false;  // ramp mode: state before the last update run kept
  std::vector<double> lastF;
  // solve scratch
  DevBuf<int32_t> solve_rve, solve_flags, order, status, steps;
  DevBuf<double> solve_F, sigma;
  DevBuf<double> omega, normal;
  // plan: per RVE a variant index per precision
  std::vector<int> plan[2];
  std::vector<size_t> lds[2];
  // tangent-linear variant per RVE (-1: the RVE fits none) and its LDS bytes
  std::vector<int> plan_tl;
  std::vector<size_t> lds_tl;
  DevBuf<double> solve_G, dsigma;
  int no_reg_variants = 0;  // planner restricted to the LDS/global-list kernels
  std::map<TuneKey, int> tuned;  // autotuned variant per bucket
  int tuned_chunks = 0;           // autotuned step chunks per solve (0: the model's)
  int plan_version = 0;          // bumped whenever the launch tables change
  // hipGraph replay of run_device_fast (mf_batch_set_graph)
  struct Graph {
    const double *F;
    int nF, update;
    double *sigma;
    int32_t *status;
    int cur, prec, version;
    static constexpr int kScratch = 10;
    const void *scratch[kScratch];  // library buffers the captured launches use
    hipGraphExec_t exec;
  };
  bool use_graph = false;
  std::vector<Graph> graphs;
  hipEvent_t ev_fork = nullptr;  // fork event of captured launches
  bool capturing = false;
  void drop_graphs() {
    if (graphs.empty()) return;
    (void)hipStreamSynchronize(stream);  // no replay still in flight
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  // one-solve-per-RVE tables kept on the device: RVE ids, launch order per
  // precision, flags per (update) and the broadcast/gathered F buffer
  DevBuf<int32_t> id_rve, id_order[2], id_flags[2];
  DevBuf<double> id_F;
  std::vector<std::pair<int, int>> id_groups[2];  // (variant, count) in order
  std::vector<size_t> id_group_lds[2];              // LDS bytes per group
  std::vector<int64_t> id_group_gs[2];              // global state per solve
  std::vector<int> id_group_chunks[2];              // step chunks per solve
  // chunk state of chunked groups (one-solve-per-RVE runs) and the per-RVE
  // chunk counters (zero between launches)
  DevBuf<float> cstate;
  DevBuf<int32_t> cflag;
  DevBuf<double> gstate;  // global-state variants' per-solve scratch
  size_t id_gtotal[2] = {0, 0};  // doubles of global state of a one-solve-per-RVE run
  int device = 0;                // HIP device the batch lives on
  DevBuf<double> lastF_dev;
  bool lastF_on_device = false;
  // energies of the last run() (FP64 energy diagnostics), 4 per RVE
  DevBuf<double> energy;
  bool have_energy = false;
  // state is the reference configuration (no update run, no set_state)
  bool at_reference = true;
  // per-RVE viscous damping (empty: visc_damp_coeff for all)
  DevBuf<double> eta_rve;
  // surrogate path buffers
  DevBuf<float> sur_desc, sur_X, sur_Y, sur_J, sur_spread;
  DevBuf<double> sur_sigma, sur_C;
  DevBuf<int32_t> sur_flag;
  ~mf_batch() {
    drop_graphs();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (int i = 0; i < kAux; ++i) {
      if (aux_done[i]) (void)hipEventDestroy(aux_done[i]);
      if (aux[i]) (void)hipStreamDestroy(aux[i]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Every fiber of the RVE uses a linear reaction with the E*A of reaction 0
// (the linear-only kernels take EA from reaction 0).
bool all_linear(const PackedRVE &p) {
  if (p.reactions.empty() || p.reactions[0].type != MF_REACTION_LINEAR) return false;
  const double ea0 = p.reactions[0].E * p.reactions[0].area();
  for (int e = 0; e < p.M; ++e) {
    const Reaction &q = p.reactions[p.rid[e]];
    if (q.type != MF_REACTION_LINEAR || q.E * q.area() != ea0) return false;
  }
  return true;
}

// At most one linear and one nonlinear reaction in the table, and a toe
// range whose expm1 argument stays within the polynomial's bound (the FP32c
// register kernel's one-linear-one-nonlinear mode).
bool mix2_ok(const PackedRVE &p) {
  if (p.reactions.empty() || p.reactions.size() > 2) return false;
  int n_nl = 0;
  for (const Reaction &q : p.reactions) {
    if (q.type != MF_REACTION_NONLINEAR) continue;
    ++n_nl;
    const double bound = 0.5 * q.B * std::max(std::fabs(q.lt * q.lt - 1), std::fabs(q.lc * q.lc - 1));
    if (!(bound <= 0.25)) return false;
  }
  return n_nl <= 1;
}

// The batch's block_size binds the kernels of precision prec when that
// precision has variants of that thread count (FP64 has none at 640 or 768
// threads, so an FP32c-only size leaves the FP64 plan to the planner).
bool block_ok(const mf_batch *b, int prec, const Variant &v) {
  const int bs = b->p.block_size;
  if (!bs || v.block == bs) return true;
  int nv = 0;
  const Variant *vs = variants(&nv);
  for (int i = 0; i < nv; ++i)
    if (vs[i].prec == prec && vs[i].block == bs && !(vs[i].flags & kTangentLinear)) return false;
  return true;
}

// Whether variant v can run RVE p (threads, registers, LDS, reactions,
// energy diagnostics) under the batch parameters.
bool variant_fits(const mf_batch *b, int prec, const Variant &v, const PackedRVE &p,
                  bool lin) {
  // tangent-linear kernels need a direction and a derivative buffer per
  // solve: only tangent-linear solve sets (make_plan_tl) run them
  if (v.flags & kTangentLinear) return false;
  if (v.flags & kGlobalState)  // FP64, any size; also the FP32c fallback
    return !(prec == MF_PREC_FP64 && b->p.energy_diagnostics) && block_ok(b, prec, v) &&
           lds_bytes(v, p.N, p.Nb, p.M, p.ell_free_entries) <= kMaxLds;
  if (v.prec != prec) return false;
  if (!block_ok(b, prec, v)) return false;
  const bool regs = (v.flags & kEllInRegs) != 0;
  const bool wave = (v.flags & kWaveRve) != 0;  // one RVE per wave: 64 threads
  const int tpr = wave ? 64 : v.block;
  if ((long)tpr * v.npt < p.Nf + (regs ? p.n_split : 0)) return false;
  if (regs && (p.max_free_deg > 2 * kLaneDegree || 2 * p.n_split > tpr ||
               p.N > (wave ? kWaveBufStride : kRegBufStride)))
    return false;
  if ((v.flags & kLinearOnly) && !lin) return false;
  if ((v.flags & kMix2) && !mix2_ok(p)) return false;
  if (((v.flags & kEnergy) != 0) != (prec == MF_PREC_FP64 && b->p.energy_diagnostics != 0))
    return false;
  if (regs && b->no_reg_variants) return false;
  const bool fpar = (v.flags & (kFiberPar | kFiberStream)) != 0;
  if (fpar && (long)v.block * variant_fpl(v.flags) < p.M) return false;
  // packed 16-bit byte offsets of 12-byte records
  if ((v.flags & kFiberStream) && 12L * p.N >= 65536) return false;
  if (prec == MF_PREC_FP32C && b->p.fp32_kernel != MF_FP32_KERNEL_AUTO &&
      (fpar != (b->p.fp32_kernel == MF_FP32_KERNEL_FIBER) ||
       wave != (b->p.fp32_kernel == MF_FP32_KERNEL_WAVE)))
    return false;
  // one RVE per wave: a workgroup holds block / 64 regions of this size
  const size_t regions = wave ? (size_t)v.block / 64 : 1;
  return regions * ((lds_bytes(v, p.N, p.Nb, p.M, p.ell_free_entries) + 15) & ~size_t(15)) <=
         kMaxLds;
}

// Autotuner bucket of an RVE: reactions, register eligibility, lanes needed
// (in wavefronts) and node count (in units of 256).
TuneKey tune_key(int prec, const PackedRVE &p, bool lin) {
  return TuneKey{prec, lin ? 1 : 0, p.max_free_deg <= 2 * kLaneDegree ? 1 : 0,
                 (p.Nf + p.n_split + 63) / 64, (p.N + 255) / 256};
}

// Picks the kernel variant of every RVE for one precision: the autotuned
// variant of the RVE's bucket when there is one, otherwise register-resident
// lists first (one node per thread before several, fewest threads), then the
// LDS/global-list kernels with the fewest threads and slots.
int make_plan(mf_batch *b, int prec) {
  int nv = 0;
  const Variant *vs = variants(&nv);
  b->plan[prec].assign(b->R, -1);
  b->lds[prec].assign(b->R, 0);
  for (int r = 0; r < b->R; ++r) {
    const PackedRVE &p = b->host[r];
    const bool lin = all_linear(p);
    int best = -1;
    auto t = b->tuned.find(tune_key(prec, p, lin));
    if (t != b->tuned.end() && !b->p.block_size && !b->p.fp32_kernel &&
        variant_fits(b, prec, vs[t->second], p, lin))
      best = t->second;
    const bool use_tuned = best >= 0;
    long best_key = 0;
    for (int i = 0; !use_tuned && i < nv; ++i) {
      const Variant &v = vs[i];
      if (!variant_fits(b, prec, v, p, lin)) continue;
      const bool regs = (v.flags & kEllInRegs) != 0;
      const bool fpar = (v.flags & kFiberPar) != 0;
      const bool fstr = (v.flags & kFiberStream) != 0;
      // reaction modes: linear-only, then one-linear-one-nonlinear, then table
      const int mode_rank = (v.flags & kLinearOnly) ? 0 : (v.flags & kMix2) ? 1 : 2;
      // Measured order.  Register-resident lists with one node per thread.
      // Then, for RVEs with nonlinear fibers: fiber-parallel with up to 3
      // fibers per thread, fiber-streamed (the shapes of up to 768 threads,
      // which hold the node lists in registers, first: the autotuner picks
      // them for 1800-3100-node toe-law RVEs), fiber-parallel with more
      // fibers, register lists with several nodes per thread; for all-linear RVEs
      // the register lists with several nodes per thread come first.  Last
      // the LDS/global-list kernels, fewest node slots per thread first (more
      // waves hide the FP64 division and square-root chains: 1000-2000 node
      // RVEs run 1.4x faster on 1024 threads x 2 slots than on 512 x 4).
      const long fpl = variant_fpl(v.flags);
      // 256 threads x 2 linear nodes per thread (three solves per CU) beats
      // one node per thread at 512+ threads: 9% on 500-node linear RVEs
      const bool reg256x2 = regs && v.block == 256 && v.npt == 2;
      // toe-law RVEs that leave more than an eighth of a 1024-thread register
      // solve idle run faster fiber-parallel (8% at 800 nodes); fuller ones
      // (1000 nodes) and the 640- and 768-thread shapes stay 2% faster in
      // registers
      const bool regbig_nl = regs && !lin && v.block == 1024 && v.npt == 1 &&
                             8L * (p.Nf + p.n_split) <= 7L * 1024;
      // one RVE per wave first wherever it holds the RVE (up to ~180
      // nodes): 6-8% faster than 256 threads per RVE on 140-180-node linear
      // RVEs (four RVEs per workgroup) and 5% on toe-law ones (one)
      const bool wave = (v.flags & kWaveRve) != 0;
      const long wave_rank = lin ? (v.block == 256 ? 0 : 10) : (v.block == 64 ? 0 : 10);
      // global-state FP64 last: only where nothing else fits.  FP64 with each
      // fiber evaluated once before the node-centric FP64 kernels for RVEs
      // with toe-law fibers (11% on 1500-node RVEs), after them for linear
      // ones, whose cheap force does not pay for the extra phase (13% slower
      // on 500-node RVEs)
      long key = (v.flags & kGlobalState) ? 2000000000L
                 : (v.flags & kFiberOnce)
                     ? (lin ? 1500000000L : 500000000L) + (long)v.npt * 10000000 + (long)v.block * 1000
                 : fpar ? (lin ? 50000000L : 15000000L) + fpl * 100000 + (long)v.block * 10 + mode_rank
                 : fstr ? (lin ? 55000000L : 15350000L) + (v.block > 768 ? 50000L : 0L) +
                              (long)v.npt * 10000 + fpl * 100 + mode_rank
                 : wave ? 5000000L + wave_rank + mode_rank
                 : reg256x2 ? 10000000L + 5000
                 : regbig_nl ? 15340000L + mode_rank
                 : regs ? (long)v.npt * 10000000L + (long)v.block * 10 + mode_rank
                        : 1000000000L + (long)v.npt * 10000000 + (long)v.block * 1000 - v.flags;
      if (best < 0 || key < best_key) {
        best = i;
        best_key = key;
      }
    }
    if (best < 0) {
      set_error("RVE " + std::to_string(r) + " (" + std::to_string(p.N) +
                " nodes) fits no kernel variant");
      return MF_ERR_UNSUPPORTED;
    }
    b->plan[prec][r] = best;
    b->lds[prec][r] = lds_bytes(vs[best], p.N, p.Nb, p.M, p.ell_free_entries);
  }
  return MF_OK;
}

// Tangent-linear variant of every RVE: the fewest threads, then the fewest
// node slots per thread, that hold its free nodes and fit the LDS.
void make_plan_tl(mf_batch *b) {
  int nv = 0;
  const Variant *vs = variants(&nv);
  b->plan_tl.assign(b->R, -1);
  b->lds_tl.assign(b->R, 0);
  for (int r = 0; r < b->R; ++r) {
    const PackedRVE &p = b->host[r];
    long best_key = 0;
    for (int i = 0; i < nv; ++i) {
      const Variant &v = vs[i];
      if (!(v.flags & kTangentLinear) || (long)v.block * v.npt < p.Nf) continue;
      const size_t L = lds_bytes(v, p.N, p.Nb, p.M, p.ell_free_entries);
      if (L > kMaxLds) continue;
      const long key = (long)v.npt * 100000 + v.block;
      if (b->plan_tl[r] < 0 || key < best_key) {
        b->plan_tl[r] = i;
        b->lds_tl[r] = L;
        best_key = key;
      }
    }
  }
}

int upload_batch(mf_batch *b) {
  const int R = b->R;
  std::vector<RveDev> rd(R);
  std::vector<double> X, mass, l0s, ell_l0;
  std::vector<int32_t> deg, sb, sD, ew, fib, fst, efr;
  std::vector<float> il, fpm;
  std::vector<ReactDev> rx;
  int64_t node_off = 0, ell_off = 0, fib_off = 0;
  int slice_off = 0;
  for (int r = 0; r < R; ++r) {
    const PackedRVE &p = b->host[r];
    RveDev &d = rd[r];
    std::memset(&d, 0, sizeof(d));
    d.N = p.N;
    d.Nf = p.Nf;
    d.Nb = p.Nb;
    d.M = p.M;
    d.node_off = node_off;
    d.ell_off = ell_off;
    d.fib_off = fib_off;
    d.slice_off = slice_off;
    d.n_slices = (int)p.slice_base.size();
    d.ell_free_entries = p.ell_free_entries;
    d.n_split = p.n_split;
    d.m_free = 0;
    for (int e = 0; e < p.M; ++e)
      if (p.fstream[4 * (size_t)e + 2] >= 0) d.m_free = e + 1;
    d.react_off = (int)rx.size();
    d.n_react = (int)p.reactions.size();
    d.c_sound = p.c_sound;
    d.scale = p.scale;
    {
      // largest power of two with max|X| * scale < 2^30
      double xm = 0;
      for (double v : p.X) xm = std::max(xm, std::fabs(v));
      int e = 0;
      std::frexp(std::max(xm, 1e-300), &e);  // xm < 2^e
      d.xq_scale = std::ldexp(1.0, 30 - e);
      d.xq_inv = (float)std::ldexp(1.0, e - 30);
    }
    X.insert(X.end(), p.X.begin(), p.X.end());
    mass.insert(mass.end(), p.mass.begin(), p.mass.end());
    deg.insert(deg.end(), p.deg.begin(), p.deg.end());
    sb.insert(sb.end(), p.slice_base.begin(), p.slice_base.end());
    sD.insert(sD.end(), p.slice_D.begin(), p.slice_D.end());
    ew.insert(ew.end(), p.ell_w.begin(), p.ell_w.end());
    ell_l0.insert(ell_l0.end(), p.ell_l0.begin(), p.ell_l0.end());
    il.insert(il.end(), p.ell_il.begin(), p.ell_il.end());
    fib.insert(fib.end(), p.fib.begin(), p.fib.end());
    fst.insert(fst.end(), p.fstream.begin(), p.fstream.end());
    efr.insert(efr.end(), p.ell_fref.begin(), p.ell_fref.end());
    fpm.insert(fpm.end(), p.fparam.begin(), p.fparam.end());
    for (const Reaction &q : p.reactions) {
      ReactDev rdv;
      make_react_dev(q, rdv);
      rx.push_back(rdv);
    }
    node_off += p.N;
    ell_off += (int64_t)p.ell_w.size();
    fib_off += p.M;
    slice_off += d.n_slices;
  }
  b->rdev = rd;
  b->total_nodes = node_off;
  hipStream_t s = b->stream;
  MF_HIP(b->rves.upload(rd, s));
  MF_HIP(b->X.upload(X, s));
  MF_HIP(b->mass.upload(mass, s));
  MF_HIP(b->deg.upload(deg, s));
  MF_HIP(b->slice_base.upload(sb, s));
  MF_HIP(b->slice_D.upload(sD, s));
  MF_HIP(b->ell_w.upload(ew, s));
  MF_HIP(b->ell_l0.upload(ell_l0, s));
  MF_HIP(b->ell_il.upload(il, s));
  MF_HIP(b->fib.upload(fib, s));
  MF_HIP(b->fstream.upload(fst, s));
  MF_HIP(b->ell_fref.upload(efr, s));
  MF_HIP(b->fparam.upload(fpm, s));
  MF_HIP(b->reacts.upload(rx, s));
  // state: x = X, u = 0, V = 1 (constructor, ...Explicit.h:115-118,250-253)
  for (int k = 0; k < 2; ++k) {
    MF_HIP(b->xs[k].upload(X, s));
    MF_HIP(b->us[k].alloc(3 * node_off));
    MF_HIP(hipMemsetAsync(b->us[k].p, 0, sizeof(double) * 3 * node_off, s));
  }
  std::vector<double> ones(R, 1.0);
  MF_HIP(b->V.upload(ones, s));
  MF_HIP(b->cur_stress.alloc(6 * (size_t)R));
  MF_HIP(b->lastF_dev.alloc(9 * (size_t)R));
  MF_HIP(hipStreamSynchronize(s));
  return MF_OK;
}

int build_identity_tables(mf_batch *b);
void fill_args(mf_batch *b, SolveArgs &a);
int launch_groups(mf_batch *b, SolveArgs a, const int32_t *order,
                  const std::vector<std::pair<int, int>> &groups,
                  const std::vector<size_t> &lds, const std::vector<int64_t> &gstride,
                  double *gstate_base = nullptr, const std::vector<int> *chunks = nullptr);
// doubles of global state per solve of a group of variant v whose largest
// RVE has N nodes (global-state variants; 0 otherwise)
int64_t group_gstride(const Variant &v, int N) {
  return (v.flags & kGlobalState) ? 12 * (int64_t)((N + 1) & ~1) : 0;
}

// Step chunks per solve for a launch group of n solves of variant v (ramp
// mode, FP32c register kernels): with S workgroups resident on the device,
// n solves of equal length take ceil(n / S) rounds, the last partly empty;
// c chunks per solve take ceil(n c / S) rounds of 1/c the length.  The
// count minimising that (plus 1% per extra chunk for the state round trip)
// is chosen, and only when n >= 2 S, so a chunk's predecessor has finished
// long before it starts.  params.chunks > 0 forces the count, else the
// autotuner's measured count when it set one.
int plan_chunks(const mf_batch *b, const Variant &v, size_t lds, int n) {
  if (b->p.mode != MF_MODE_RAMP || v.prec != MF_PREC_FP32C || !(v.flags & kEllInRegs) ||
      (v.flags & kWaveRve))
    return 1;
  const int force = b->p.chunks > 0 ? b->p.chunks : b->tuned_chunks;
  if (force > 0) return std::max(1, std::min(force, b->p.n_steps));
  const int occ = chunk_occupancy(v, lds);
  int cus = 0;
  if (occ <= 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) !=
                      hipSuccess || cus <= 0)
    return 1;
  const long S = (long)occ * cus;
  if (n < 2 * S) return 1;
  int best = 1;
  double best_t = 1e300;
  for (int c = 1; c <= 4 && c <= b->p.n_steps; ++c) {
    const double t = (double)((n * (long)c + S - 1) / S) / c * (1.0 + 0.01 * (c - 1));
    if (t < best_t - 1e-12) {
      best_t = t;
      best = c;
    }
  }
  return best;
}

int create_from_networks(std::vector<Network> &nets, const mf_solver_params *params,
                         mf_batch **out) {
  if (!out) {
    set_error("out is NULL");
    return MF_ERR_INVALID;
  }
  *out = nullptr;
  mf_solver_params p;
  if (params)
    p = *params;
  else
    mf_default_params(&p);
  if (!params_valid(p)) {
    set_error("invalid mf_solver_params");
    return MF_ERR_INVALID;
  }
  std::unique_ptr<mf_batch> b(new mf_batch());
  b->p = p;
  b->R = (int)nets.size();
  b->policy = p.reaction_policy;
  b->host.resize(b->R);
  std::vector<int> rc(b->R, MF_OK);
  std::vector<std::string> errs(b->R);
  parallel_for(b->R, [&](int r) {
    rc[r] = pack_rve(nets[r], p.reaction_policy, b->host[r]);
  });
  for (int r = 0; r < b->R; ++r)
    if (rc[r] != MF_OK) {
      set_error("RVE " + std::to_string(r) + ": invalid network");
      return rc[r];
    }
  for (int prec = 0; prec < 2; ++prec) {
    int e = make_plan(b.get(), prec);
    if (e != MF_OK) return e;
  }
  make_plan_tl(b.get());
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device available");
    return MF_ERR_HIP;
  }
  MF_HIP(hipGetDevice(&b->device));
  MF_HIP(hipStreamCreate(&b->stream));
  MF_HIP(hipEventCreate(&b->ev0));
  MF_HIP(hipEventCreate(&b->ev1));
  MF_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  for (int i = 0; i < mf_batch::kAux; ++i) {
    MF_HIP(hipStreamCreate(&b->aux[i]));
    MF_HIP(hipEventCreateWithFlags(&b->aux_done[i], hipEventDisableTiming));
  }
  int e = upload_batch(b.get());
  if (e != MF_OK) return e;
  e = build_identity_tables(b.get());
  if (e != MF_OK) return e;
  *out = b.release();
  return MF_OK;
}

// Launch order of n solves (solve i runs RVE rve(i)): grouped by kernel
// variant, groups in decreasing order of their largest solve (fibers), and
// largest solves first inside a group.  Longest workgroups start first, so
// they do not trail the batch; a small group of a wider variant also gets
// empty CUs before the big groups fill them (a CU running the common variant
// may have no room left for its wider workgroups).  One-RVE-per-wave groups
// go first whatever their fiber counts: a whole solve on one wave (up to
// three nodes per lane) takes longer than the larger RVEs' workgroups, and
// launched last the group trailed a 16384-RVE batch by 12 ms.
template <class RveOf>
std::vector<int32_t> launch_order(const mf_batch *b, int prec, int n, RveOf rve) {
  int nv = 0;
  const Variant *vs = variants(&nv);
  std::map<int, int> max_m;  // variant -> largest fiber count
  for (int i = 0; i < n; ++i) {
    const int r = rve(i), v = b->plan[prec][r];
    max_m[v] = std::max(max_m[v], b->host[r].M);
  }
  auto per_wave = [&](int v) { return v >= 0 && v < nv && (vs[v].flags & kWaveRve) != 0; };
  std::vector<int32_t> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const int ra = rve(a), rc = rve(c);
    const int va = b->plan[prec][ra], vc = b->plan[prec][rc];
    if (va != vc) {
      const bool wa = per_wave(va), wc = per_wave(vc);
      if (wa != wc) return wa;
      const int ma = max_m.at(va), mc = max_m.at(vc);
      return ma != mc ? ma > mc : va < vc;
    }
    return b->host[ra].M > b->host[rc].M;
  });
  return order;
}

// Launches one set of solves.  solve_* arrays live on the host and are
// uploaded; sigma/status/steps land in the device scratch.
struct SolveSet {
  std::vector<int32_t> rve, flags;
  std::vector<double> F;  // 9 per solve
  int prec = MF_PREC_FP64;
  // tangent-linear solves: direction of each perturbation (9 per solve);
  // d sigma lands in b->dsigma
  bool tl = false;
  std::vector<double> G;
};

int launch_set(mf_batch *b, const SolveSet &S, const double *x_in,
               const double *u_in, double *x_out, double *u_out,
               double *sigma_dev = nullptr, int32_t *status_dev = nullptr) {
  const int n = (int)S.rve.size();
  hipStream_t s = b->stream;
  MF_HIP(b->solve_rve.ensure(n));
  MF_HIP(b->solve_flags.ensure(n));
  MF_HIP(b->solve_F.ensure(9 * (size_t)n));
  MF_HIP(b->order.ensure(n));
  MF_HIP(b->sigma.ensure(6 * (size_t)n));
  MF_HIP(b->status.ensure(n));
  MF_HIP(b->steps.ensure(n));
  const int prec = S.prec;
  if (S.tl) {
    for (int i = 0; i < n; ++i)
      if (b->plan_tl[S.rve[i]] < 0) {
        set_error("RVE " + std::to_string(S.rve[i]) + " (" + std::to_string(b->host[S.rve[i]].N) +
                  " nodes) is too large for the tangent-linear kernel; use the central scheme");
        return MF_ERR_UNSUPPORTED;
      }
    MF_HIP(b->solve_G.ensure(9 * (size_t)n));
    MF_HIP(b->dsigma.ensure(6 * (size_t)n));
    MF_HIP(hipMemcpyAsync(b->solve_G.p, S.G.data(), sizeof(double) * 9 * n,
                          hipMemcpyHostToDevice, s));
  }
  const std::vector<int> &plan = S.tl ? b->plan_tl : b->plan[prec];
  const std::vector<size_t> &plds = S.tl ? b->lds_tl : b->lds[prec];
  std::vector<int32_t> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  if (S.tl)  // grouped by variant, largest RVEs first
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const int vx = plan[S.rve[x]], vy = plan[S.rve[y]];
      if (vx != vy) return vx > vy;
      return b->host[S.rve[x]].M > b->host[S.rve[y]].M;
    });
  else
    order = launch_order(b, prec, n, [&](int i) { return S.rve[i]; });
  MF_HIP(hipMemcpyAsync(b->solve_rve.p, S.rve.data(), sizeof(int32_t) * n,
                        hipMemcpyHostToDevice, s));
  MF_HIP(hipMemcpyAsync(b->solve_flags.p, S.flags.data(), sizeof(int32_t) * n,
                        hipMemcpyHostToDevice, s));
  MF_HIP(hipMemcpyAsync(b->solve_F.p, S.F.data(), sizeof(double) * 9 * n,
                        hipMemcpyHostToDevice, s));
  MF_HIP(hipMemcpyAsync(b->order.p, order.data(), sizeof(int32_t) * n,
                        hipMemcpyHostToDevice, s));
  SolveArgs a;
  std::memset(&a, 0, sizeof(a));
  a.rves = b->rves.p;
  a.X = b->X.p;
  a.mass = b->mass.p;
  a.deg = b->deg.p;
  a.slice_base = b->slice_base.p;
  a.slice_D = b->slice_D.p;
  a.ell_w = b->ell_w.p;
  a.ell_l0 = b->ell_l0.p;
  a.ell_il = b->ell_il.p;
  a.fib = b->fib.p;
  a.fstream = b->fstream.p;
  a.ell_fref = b->ell_fref.p;
  a.fparam = b->fparam.p;
  a.reacts = b->reacts.p;
  a.x_in = x_in;
  a.u_in = u_in;
  a.x_out = x_out;
  a.u_out = u_out;
  a.V = b->V.p;
  a.solve_rve = b->solve_rve.p;
  a.solve_F = b->solve_F.p;
  a.solve_flags = b->solve_flags.p;
  a.sigma = sigma_dev ? sigma_dev : b->sigma.p;
  a.status = status_dev ? status_dev : b->status.p;
  a.steps = b->steps.p;
  a.mode = b->p.mode;
  a.n_steps = b->p.n_steps;
  a.max_steps = b->p.max_steps;
  a.freq = b->p.dt_update_freq;
  a.eta = b->p.visc_damp_coeff;
  a.eta_rve = b->eta_rve.p;
  a.amplitude = b->p.amplitude;
  a.energy_eps = b->p.energy_check_eps;
  a.crit = b->p.crit_time_scale_factor;
  a.tol = b->p.tolerance;
  if (S.tl) {
    a.solve_G = b->solve_G.p;
    a.dsigma = b->dsigma.p;
  }
  std::vector<std::pair<int, int>> groups;
  std::vector<size_t> glds;
  std::vector<int64_t> ggs;
  int nv = 0;
  const Variant *vs = variants(&nv);
  for (int i = 0; i < n;) {
    int v = plan[S.rve[order[i]]];
    int j = i;
    size_t L = 0;
    int maxN = 0;
    while (j < n && plan[S.rve[order[j]]] == v) {
      L = std::max(L, plds[S.rve[order[j]]]);
      maxN = std::max(maxN, b->host[S.rve[order[j]]].N);
      ++j;
    }
    groups.push_back({v, j - i});
    glds.push_back(L);
    ggs.push_back(group_gstride(vs[v], maxN));
    i = j;
  }
  return launch_groups(b, a, b->order.p, groups, glds, ggs);
}

void fill_args(mf_batch *b, SolveArgs &a) {
  std::memset(&a, 0, sizeof(a));
  a.rves = b->rves.p;
  a.X = b->X.p;
  a.mass = b->mass.p;
  a.deg = b->deg.p;
  a.slice_base = b->slice_base.p;
  a.slice_D = b->slice_D.p;
  a.ell_w = b->ell_w.p;
  a.ell_l0 = b->ell_l0.p;
  a.ell_il = b->ell_il.p;
  a.fib = b->fib.p;
  a.fstream = b->fstream.p;
  a.ell_fref = b->ell_fref.p;
  a.fparam = b->fparam.p;
  a.reacts = b->reacts.p;
  a.V = b->V.p;
  a.steps = b->steps.p;
  a.mode = b->p.mode;
  a.n_steps = b->p.n_steps;
  a.max_steps = b->p.max_steps;
  a.freq = b->p.dt_update_freq;
  a.eta = b->p.visc_damp_coeff;
  a.eta_rve = b->eta_rve.p;
  a.amplitude = b->p.amplitude;
  a.energy_eps = b->p.energy_check_eps;
  a.crit = b->p.crit_time_scale_factor;
  a.tol = b->p.tolerance;
}

// Device-resident tables for the common case of one solve per RVE.
int build_identity_tables(mf_batch *b) {
  const int R = b->R;
  b->plan_version++;
  b->drop_graphs();
  hipStream_t s = b->stream;
  std::vector<int32_t> ids(R);
  for (int r = 0; r < R; ++r) ids[r] = r;
  MF_HIP(b->id_rve.upload(ids, s));
  MF_HIP(b->id_F.alloc(9 * (size_t)R));
  MF_HIP(b->steps.ensure(R));
  size_t ctotal = 0;  // chunk state of both precisions' plans (the larger is kept)
  size_t cneed = 0;
  for (int prec = 0; prec < 2; ++prec) {
    ctotal = 0;
    const std::vector<int32_t> order = launch_order(b, prec, R, [](int i) { return i; });
    b->id_groups[prec].clear();
    b->id_group_lds[prec].clear();
    b->id_group_gs[prec].clear();
    int nv = 0;
    const Variant *vs = variants(&nv);
    b->id_group_chunks[prec].clear();
    size_t gtotal = 0;
    for (int i = 0; i < R;) {
      int v = b->plan[prec][order[i]], j = i;
      size_t L = 0;
      int maxN = 0;
      while (j < R && b->plan[prec][order[j]] == v) {
        L = std::max(L, b->lds[prec][order[j]]);
        maxN = std::max(maxN, b->host[order[j]].N);
        ++j;
      }
      b->id_groups[prec].push_back({v, j - i});
      b->id_group_lds[prec].push_back(L);
      b->id_group_gs[prec].push_back(group_gstride(vs[v], maxN));
      gtotal += (size_t)(j - i) * (size_t)b->id_group_gs[prec].back();
      const int c = plan_chunks(b, vs[v], L, j - i);
      b->id_group_chunks[prec].push_back(c);
      if (c > 1) ctotal += (size_t)(j - i) * (size_t)chunk_stride(vs[v].block, vs[v].npt);
      i = j;
    }
    // allocated here, so a captured run launch never allocates
    b->id_gtotal[prec] = gtotal;
    MF_HIP(b->gstate.ensure(gtotal));
    cneed = std::max(cneed, ctotal);
    MF_HIP(b->id_order[prec].upload(order, s));
    // flags: [0..R) update=false, [R..2R) update=true
    std::vector<int32_t> fl(2 * (size_t)R);
    const bool ramp = b->p.mode == MF_MODE_RAMP;
    for (int r = 0; r < R; ++r) {
      fl[r] = ramp ? 0 : kWriteX;
      fl[R + r] = kWriteX | kWriteU;
    }
    MF_HIP(b->id_flags[prec].upload(fl, s));
  }
  if (cneed > 0) {
    MF_HIP(b->cstate.ensure(cneed));
    if (b->cflag.n < (size_t)R) {
      MF_HIP(b->cflag.alloc(R));
      MF_HIP(hipMemsetAsync(b->cflag.p, 0, sizeof(int32_t) * R, s));
    }
  }
  MF_HIP(hipStreamSynchronize(s));
  return MF_OK;
}

// Launches groups (variant, solve count, LDS bytes) of consecutive entries
// of `order`; group 0 runs on the batch stream, the others on the auxiliary
// streams, all after ev0 and joined back before ev1.  gstate_base, when
// given, holds the groups' global state (the caller sized it); otherwise the
// batch's own buffer is used and grown as needed.
int launch_groups(mf_batch *b, SolveArgs a, const int32_t *order,
                  const std::vector<std::pair<int, int>> &groups,
                  const std::vector<size_t> &lds, const std::vector<int64_t> &gstride,
                  double *gstate_base, const std::vector<int> *chunks) {
  int nv = 0;
  const Variant *vs = variants(&nv);
  hipStream_t s = b->stream;
  size_t gtotal = 0;
  for (size_t gi = 0; gi < groups.size(); ++gi)
    gtotal += (size_t)groups[gi].second * (size_t)gstride[gi];
  if (!gstate_base && gtotal > b->gstate.n) {
    if (b->capturing) {
      set_error("global-state scratch would grow inside a captured launch");
      return MF_ERR_STATE;
    }
    MF_HIP(b->gstate.ensure(gtotal));
  }
  size_t goff = 0, coff = 0;
  // under capture the fork uses its own event and the timing events stay
  // outside the graph
  hipEvent_t fork = b->capturing ? b->ev_fork : b->ev0;
  MF_HIP(hipEventRecord(fork, s));
  bool used[mf_batch::kAux] = {false, false, false};
  int off = 0;
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    hipStream_t st = s;
    if (gi > 0) {
      const int k = (int)((gi - 1) % mf_batch::kAux);
      st = b->aux[k];
      if (!used[k]) MF_HIP(hipStreamWaitEvent(st, fork, 0));
      used[k] = true;
    }
    a.order = order + off;
    a.gstate = gstride[gi] ? (gstate_base ? gstate_base : b->gstate.p) + goff : nullptr;
    a.gstride = gstride[gi];
    goff += (size_t)groups[gi].second * (size_t)gstride[gi];
    const Variant &vg = vs[groups[gi].first];
    const int nc = chunks ? (*chunks)[gi] : 1;
    a.n_chunks = nc;
    a.group_n = groups[gi].second;
    a.chunk_steps = nc > 1 ? (b->p.n_steps + nc - 1) / nc : 0;
    a.cstride = nc > 1 ? chunk_stride(vg.block, vg.npt) : 0;
    a.cstate = nc > 1 ? b->cstate.p + coff : nullptr;
    a.cflag = nc > 1 ? b->cflag.p : nullptr;
    if (nc > 1) {
      coff += (size_t)groups[gi].second * (size_t)a.cstride;
      if (coff > b->cstate.n || !b->cflag.p) {
        set_error("step-chunk state not allocated for this launch");
        return MF_ERR_STATE;
      }
    }
    hipError_t e = launch_solve(vg, a, groups[gi].second * nc, lds[gi], st);
    if (e != hipSuccess) {
      set_error(std::string("solve launch: ") + hipGetErrorString(e));
      return MF_ERR_HIP;
    }
    off += groups[gi].second;
  }
  for (int k = 0; k < mf_batch::kAux; ++k)
    if (used[k]) {
      MF_HIP(hipEventRecord(b->aux_done[k], b->aux[k]));
      MF_HIP(hipStreamWaitEvent(s, b->aux_done[k], 0));
    }
  if (!b->capturing) MF_HIP(hipEventRecord(b->ev1, s));
  b->last_launches = (int)groups.size();
  return MF_OK;
}

int finish_timing(mf_batch *b) {
  MF_HIP(hipEventSynchronize(b->ev1));
  float ms = 0;
  MF_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  return MF_OK;
}

int check_F(mf_batch *b, int nF, int update) {
  if (nF != 1 && nF != b->R) {
    set_error("F must hold 1 or n_rve deformation gradients");
    return MF_ERR_INVALID;
  }
  // updateVolume requires one F per RVE (...Base.h:136-142)
  if (update && nF != b->R) {
    set_error("The volume and deformation gradients must have the same extent "
              "corresponding to the number of RVEs.");
    return MF_ERR_INVALID;
  }
  return MF_OK;
}

// Device scratch of the allocation-free run: F staging for a broadcast F,
// per-solve step counters and statuses, global state.
struct RunScratch {
  double *F_stage;
  int32_t *steps;
  int32_t *status;
  double *gstate;
};

// run() with F on the device and one solve per RVE: no host round trip.
// With `na` every buffer the launches write besides the caller's comes from
// it or was allocated with the batch, and nothing is allocated.
int run_device_fast(mf_batch *b, const double *F_dev, int nF, int update,
                    double *sigma_dev, int32_t *status_dev,
                    const RunScratch *na = nullptr) {
  const int R = b->R;
  hipStream_t s = b->stream;
  const bool ramp = b->p.mode == MF_MODE_RAMP;
  const int prec = b->p.precision;
  if (update) {
    hipError_t e = launch_update_volume(F_dev, nF, b->V.p, R, s);
    if (e != hipSuccess) {
      set_error(hipGetErrorString(e));
      return MF_ERR_HIP;
    }
  }
  const double *Fs = F_dev;
  if (nF == 1) {
    double *Fb = na ? na->F_stage : b->id_F.p;
    hipError_t e = launch_broadcast_F(F_dev, Fb, R, s);
    if (e != hipSuccess) {
      set_error(hipGetErrorString(e));
      return MF_ERR_HIP;
    }
    Fs = Fb;
  }
  if (!na) {
    MF_HIP(b->sigma.ensure(6 * (size_t)R));
    MF_HIP(b->status.ensure(R));
  }
  SolveArgs a;
  fill_args(b, a);
  if (na) a.steps = na->steps;
  const int c = b->cur;
  a.x_in = b->xs[c].p;
  a.u_in = b->us[c].p;
  a.x_out = ramp ? b->xs[c ^ 1].p : b->xs[c].p;
  a.u_out = ramp ? b->us[c ^ 1].p : b->us[c].p;
  a.solve_rve = b->id_rve.p;
  a.solve_F = Fs;
  a.solve_flags = b->id_flags[prec].p + (update ? R : 0);
  a.sigma = sigma_dev ? sigma_dev : b->sigma.p;
  a.status = status_dev ? status_dev : na ? na->status : b->status.p;
  b->have_energy = prec == MF_PREC_FP64 && b->p.energy_diagnostics;
  if (b->have_energy) {
    if (na) {
      set_error("energy diagnostics need a library buffer; the allocation-free run "
                "does not track them");
      return MF_ERR_UNSUPPORTED;
    }
    MF_HIP(b->energy.ensure(4 * (size_t)R));
    a.energy = b->energy.p;
  }
  {
    // the allocation-free run has no chunk state: one piece per solve
    int e = launch_groups(b, a, b->id_order[prec].p, b->id_groups[prec],
                          b->id_group_lds[prec], b->id_group_gs[prec],
                          na ? na->gstate : nullptr, na ? nullptr : &b->id_group_chunks[prec]);
    if (e != MF_OK) return e;
  }
  if (!ramp) b->at_reference = false;  // relax runs keep the trial x (Q5)
  if (update) {
    MF_HIP(hipMemcpyAsync(b->cur_stress.p, a.sigma, sizeof(double) * 6 * R,
                          hipMemcpyDeviceToDevice, s));
    b->have_run = true;
    b->at_reference = false;
    if (ramp) {
      b->cur ^= 1;
      b->have_prev = true;
      if (!na) MF_HIP(b->lastF_dev.ensure(9 * (size_t)R));  // allocated with the batch
      MF_HIP(hipMemcpyAsync(b->lastF_dev.p, Fs, sizeof(double) * 9 * R,
                            hipMemcpyDeviceToDevice, s));
      b->lastF_on_device = true;
    }
  }
  return MF_OK;
}

// run_device_fast through a captured hipGraph: the first call with a given
// (buffers, flags, state parity, launch tables) captures the launch sequence
// (volume update, F broadcast, all launch groups on the batch and auxiliary
// streams, commits), later calls replay it.  Host bookkeeping runs on every
// call; device buffers are allocated before capture.
int run_device_graph(mf_batch *b, const double *F_dev, int nF, int update, double *sigma_dev,
                     int32_t *status_dev) {
  const int R = b->R;
  hipStream_t s = b->stream;
  const int prec = b->p.precision;
  const bool ramp = b->p.mode == MF_MODE_RAMP;
  MF_HIP(b->sigma.ensure(6 * (size_t)R));
  MF_HIP(b->status.ensure(R));
  if (prec == MF_PREC_FP64 && b->p.energy_diagnostics) MF_HIP(b->energy.ensure(4 * (size_t)R));
  if (update && ramp) MF_HIP(b->lastF_dev.ensure(9 * (size_t)R));
  // every library buffer the captured launches read or write: a buffer that
  // was reallocated since the capture makes the graph stale
  const void *scratch[mf_batch::Graph::kScratch] = {
      b->sigma.p, b->status.p, b->energy.p, b->lastF_dev.p,
      b->eta_rve.p, b->steps.p, b->gstate.p, b->cur_stress.p, b->cstate.p, b->cflag.p};
  mf_batch::Graph *g = nullptr;
  for (auto &x : b->graphs)
    if (x.F == F_dev && x.nF == nF && x.update == update && x.sigma == sigma_dev &&
        x.status == status_dev && x.cur == b->cur && x.prec == prec &&
        x.version == b->plan_version && std::equal(scratch, scratch + mf_batch::Graph::kScratch, x.scratch))
      g = &x;
  if (!g) {
    if (b->graphs.size() >= 8) b->drop_graphs();
    // capture with the state as it is now; the bookkeeping is redone below
    const int cur0 = b->cur;
    const bool hr = b->have_run, hp = b->have_prev, ar = b->at_reference,
               lf = b->lastF_on_device, he = b->have_energy;
    hipGraph_t graph = nullptr;
    MF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    b->capturing = true;
    int e = run_device_fast(b, F_dev, nF, update, sigma_dev, status_dev);
    b->capturing = false;
    hipError_t he2 = hipStreamEndCapture(s, &graph);
    b->cur = cur0;
    b->have_run = hr;
    b->have_prev = hp;
    b->at_reference = ar;
    b->lastF_on_device = lf;
    b->have_energy = he;
    if (e != MF_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return e;
    }
    MF_HIP(he2);
    hipGraphExec_t exec = nullptr;
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MF_HIP(ie);
    mf_batch::Graph ng{F_dev, nF, update, sigma_dev, status_dev, b->cur, prec, b->plan_version,
                       {}, exec};
    std::copy(scratch, scratch + mf_batch::Graph::kScratch, ng.scratch);
    b->graphs.push_back(ng);
    g = &b->graphs.back();
  }
  MF_HIP(hipEventRecord(b->ev0, s));
  MF_HIP(hipGraphLaunch(g->exec, s));
  MF_HIP(hipEventRecord(b->ev1, s));
  // host bookkeeping of run_device_fast
  b->have_energy = prec == MF_PREC_FP64 && b->p.energy_diagnostics;
  if (!ramp) b->at_reference = false;
  if (update) {
    b->have_run = true;
    b->at_reference = false;
    if (ramp) {
      b->cur ^= 1;
      b->have_prev = true;
      b->lastF_on_device = true;
    }
  }
  return MF_OK;
}

int copy_out(mf_batch *b, double *sigma, int32_t *status, int n) {
  hipStream_t s = b->stream;
  if (sigma)
    MF_HIP(hipMemcpyAsync(sigma, b->sigma.p, sizeof(double) * 6 * n,
                          hipMemcpyDeviceToHost, s));
  if (status)
    MF_HIP(hipMemcpyAsync(status, b->status.p, sizeof(int32_t) * n,
                          hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  return MF_OK;
}

// Tangent: relax mode follows the reference sequence (one trial launch per
// column, each trial starting from the previous trial's coordinates);
// ramp mode runs every perturbed solve from the state before the last
// update run in one launch.  C is left on the device in `Cdev`.
int tangent_impl(mf_batch *b, double *Cdev, std::vector<int32_t> *status_out) {
  const int R = b->R;
  hipStream_t s = b->stream;
  const double h = b->p.fd_step;
  const bool central = b->p.tangent_scheme == MF_TANGENT_CENTRAL;
  if (!b->have_run) {
    set_error("computeMaterialStiffness needs a previous run with update_coords");
    return MF_ERR_STATE;
  }
  // The forward scheme differences each perturbed solve against a base
  // stress of the same precision.  When the perturbed solves run at another
  // precision than run(), the committed stress cannot serve as that base
  // (their difference would carry the precision gap divided by h), so an
  // unperturbed solve at tangent_precision provides it.
  const bool own_base = !central && b->p.tangent_precision != b->p.precision;
  DevBuf<double> sig_p;
  MF_HIP(sig_p.alloc(6 * (size_t)R));
  if (b->p.mode == MF_MODE_RELAX) {
    DevBuf<double> sig_m, sig_b;
    if (central) MF_HIP(sig_m.alloc(6 * (size_t)R));
    std::vector<int32_t> st_all(R, MF_STATUS_OK), st(R);
    // one trial: every RVE from the current coordinates (kept afterwards,
    // as the reference's trials keep theirs), u restored
    auto trial = [&](const double Fp[9], double *sig) -> int {
      SolveSet S;
      S.prec = b->p.tangent_precision;
      S.rve.resize(R);
      S.flags.assign(R, kWriteX);
      S.F.resize(9 * (size_t)R);
      for (int r = 0; r < R; ++r) {
        S.rve[r] = r;
        std::copy(Fp, Fp + 9, S.F.begin() + 9 * (size_t)r);
      }
      const int c = b->cur;
      int e = launch_set(b, S, b->xs[c].p, b->us[c].p, b->xs[c].p, b->us[c].p, sig);
      if (e != MF_OK) return e;
      MF_HIP(hipMemcpyAsync(st.data(), b->status.p, sizeof(int32_t) * R,
                            hipMemcpyDeviceToHost, s));
      MF_HIP(hipStreamSynchronize(s));
      for (int r = 0; r < R; ++r)
        if (st_all[r] == MF_STATUS_OK) st_all[r] = st[r];
      return MF_OK;
    };
    const double* base = b->cur_stress.p;
    if (own_base) {
      // zero-increment trial: the committed state relaxed at tangent_precision
      MF_HIP(sig_b.alloc(6 * (size_t)R));
      const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      int e = trial(I3, sig_b.p);
      if (e != MF_OK) return e;
      base = sig_b.p;
    }
    for (int i = 0; i < 6; ++i) {
      for (int sign = 0; sign < (central ? 2 : 1); ++sign) {
        double Fp[9];
        perturbation(i, sign ? -h : h, Fp);
        int e = trial(Fp, sign ? sig_m.p : sig_p.p);
        if (e != MF_OK) return e;
      }
      hipError_t e = central
          ? launch_fd(sig_p.p, sig_m.p, Cdev, R, i, 2 * h, 6, 6, s)
          : launch_fd(sig_p.p, base, Cdev, R, i, h, 6, 6, s);
      if (e != hipSuccess) {
        set_error(hipGetErrorString(e));
        return MF_ERR_HIP;
      }
    }
    MF_HIP(hipStreamSynchronize(s));
    if (status_out) *status_out = st_all;
    return MF_OK;
  }
  // ramp mode
  if (!b->have_prev) {
    set_error("ramp-mode tangent needs the state of the last update run");
    return MF_ERR_STATE;
  }
  if (b->lastF_on_device) {
    b->lastF.assign(9 * (size_t)R, 0.0);
    MF_HIP(hipMemcpyAsync(b->lastF.data(), b->lastF_dev.p, sizeof(double) * 9 * R,
                          hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    b->lastF_on_device = false;
  }
  if (b->p.tangent_scheme == MF_TANGENT_LINEAR) {
    // 6 tangent-linear solves per RVE from the state before the last update
    // run: F_last with its derivative along E_i F_last
    SolveSet S;
    S.prec = MF_PREC_FP32C;
    S.tl = true;
    for (int r = 0; r < R; ++r) {
      const double *F0 = b->lastF.data() + 9 * (size_t)r;
      for (int i = 0; i < 6; ++i) {
        double E[9], G[9];
        perturbation_rate(i, E);
        matmul3(E, F0, G);
        S.rve.push_back(r);
        S.flags.push_back(0);
        S.F.insert(S.F.end(), F0, F0 + 9);
        S.G.insert(S.G.end(), G, G + 9);
      }
    }
    const int p = b->cur ^ 1;
    int e = launch_set(b, S, b->xs[p].p, b->us[p].p, nullptr, nullptr);
    if (e != MF_OK) return e;
    hipError_t he = launch_tl_assemble(b->dsigma.p, Cdev, R, s);
    if (he != hipSuccess) {
      set_error(hipGetErrorString(he));
      return MF_ERR_HIP;
    }
    if (status_out) {
      std::vector<int32_t> st(S.rve.size());
      MF_HIP(hipMemcpyAsync(st.data(), b->status.p, sizeof(int32_t) * st.size(),
                            hipMemcpyDeviceToHost, s));
      MF_HIP(hipStreamSynchronize(s));
      status_out->assign(R, MF_STATUS_OK);
      for (size_t k = 0; k < st.size(); ++k)
        if ((*status_out)[S.rve[k]] == MF_STATUS_OK) (*status_out)[S.rve[k]] = st[k];
    }
    MF_HIP(hipStreamSynchronize(s));
    return MF_OK;
  }
  const int per = central ? 13 : 7;
  SolveSet S;
  S.prec = b->p.tangent_precision;
  for (int r = 0; r < R; ++r) {
    const double *F0 = b->lastF.data() + 9 * (size_t)r;
    if (own_base) {  // the update run's solve again, at tangent_precision
      S.rve.push_back(r);
      S.flags.push_back(0);
      S.F.insert(S.F.end(), F0, F0 + 9);
    }
    for (int k = 1; k < per; ++k) {
      int i = central ? (k - 1) / 2 : k - 1;
      double sgn = (central && ((k - 1) % 2 == 1)) ? -1.0 : 1.0;
      double Fp[9], Ft[9];
      perturbation(i, sgn * h, Fp);
      matmul3(Fp, F0, Ft);
      S.rve.push_back(r);
      S.flags.push_back(0);
      S.F.insert(S.F.end(), Ft, Ft + 9);
    }
  }
  const int p = b->cur ^ 1;  // state before the last update run
  int e = launch_set(b, S, b->xs[p].p, b->us[p].p, nullptr, nullptr);
  if (e != MF_OK) return e;
  // [base, perturbed...] per RVE for the FD kernel: the solve set's own
  // layout when it holds the base, else the committed stress in front
  DevBuf<double> all;
  const double *rows = b->sigma.p;
  if (!own_base) {
    MF_HIP(all.alloc((size_t)R * per * 6));
    MF_HIP(hipMemcpy2DAsync(all.p, sizeof(double) * 6 * per, b->cur_stress.p,
                            sizeof(double) * 6, sizeof(double) * 6, R,
                            hipMemcpyDeviceToDevice, s));
    MF_HIP(hipMemcpy2DAsync(all.p + 6, sizeof(double) * 6 * per, b->sigma.p,
                            sizeof(double) * 6 * (per - 1),
                            sizeof(double) * 6 * (per - 1), R,
                            hipMemcpyDeviceToDevice, s));
    rows = all.p;
  }
  hipError_t he = launch_fd_all(rows, Cdev, R,
                                central ? MF_TANGENT_CENTRAL : MF_TANGENT_FORWARD,
                                h, s);
  if (he != hipSuccess) {
    set_error(hipGetErrorString(he));
    return MF_ERR_HIP;
  }
  if (status_out) {
    std::vector<int32_t> st(S.rve.size());
    MF_HIP(hipMemcpyAsync(st.data(), b->status.p, sizeof(int32_t) * st.size(),
                          hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    status_out->assign(R, MF_STATUS_OK);
    for (size_t k = 0; k < st.size(); ++k)
      if ((*status_out)[S.rve[k]] == MF_STATUS_OK) (*status_out)[S.rve[k]] = st[k];
  }
  MF_HIP(hipStreamSynchronize(s));
  return MF_OK;
}

// ---- scratch of the allocation-free run ------------------------------------
// Device memory of mf_batch_run_device_noalloc, held by every GPU the library
// runs on until the process ends.  Host code never uses the array's name as
// an address: each device's copy is located once with hipGetSymbolAddress and
// that pointer is what kernels and copies receive.
static __device__ __align__(256) unsigned char scratch[MF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on the current device (looked up on first use).
int scratch_address(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index out of range");
    return MF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return MF_ERR_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
  return MF_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" {

const char *mf_last_error(void) { return g_last_error.c_str(); }
int mf_version(void) { return 1; }

void mf_default_params(mf_solver_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->mode = MF_MODE_RELAX;
  p->precision = MF_PREC_FP64;
  p->reaction_policy = MF_REACT_BATCHED_REFERENCE;
  p->tangent_scheme = MF_TANGENT_FORWARD;
  p->tangent_precision = MF_PREC_FP64;
  p->n_steps = 5000;
  p->max_steps = 2000000;
  p->dt_update_freq = 100;
  p->block_size = 0;
  p->visc_damp_coeff = 100.0;
  p->crit_time_scale_factor = 0.8;
  p->tolerance = 1e-6;
  p->fd_step = 1e-5;
}

void mf_default_gen_params(mf_gen_params *g) {
  std::memset(g, 0, sizeof(*g));
  g->seed = 1;
  g->n_nodes = 500;
  g->k_neighbors = 4;
  g->boundary_band = 0.02;
  g->radius = 3.5e-8;
  g->E = 4.32e7;
  g->density = 1000.0;
  g->nonlinear_fraction = 0.0;
  g->B = 1.2;
  g->length_ratio_trns = 1.05;
}

int mf_batch_create(const mf_network *nets, int32_t n_rve,
                    const mf_solver_params *params, mf_batch **out) {
  if (!nets || n_rve <= 0) {
    set_error("mf_batch_create: no networks");
    return MF_ERR_INVALID;
  }
  std::vector<Network> v(n_rve);
  for (int r = 0; r < n_rve; ++r) {
    int e = network_from_desc(nets[r], v[r]);
    if (e != MF_OK) return e;
  }
  return create_from_networks(v, params, out);
}

int mf_batch_create_generated(const mf_gen_params *g, int32_t n_rve,
                              const int32_t *n_nodes,
                              const mf_solver_params *params, mf_batch **out) {
  if (!g || n_rve <= 0) {
    set_error("mf_batch_create_generated: bad arguments");
    return MF_ERR_INVALID;
  }
  std::vector<Network> v(n_rve);
  std::vector<int> rc(n_rve, MF_OK);
  parallel_for(n_rve, [&](int r) {
    mf_gen_params q = *g;
    q.seed = g->seed + (uint64_t)r;
    if (n_nodes) q.n_nodes = n_nodes[r];
    rc[r] = generate_network(q, v[r]);
  });
  for (int r = 0; r < n_rve; ++r)
    if (rc[r] != MF_OK) return rc[r];
  return create_from_networks(v, params, out);
}

void mf_batch_destroy(mf_batch *b) { delete b; }
int mf_batch_num_rves(const mf_batch *b) { return b ? b->R : 0; }

int mf_batch_set_params(mf_batch *b, const mf_solver_params *params) {
  if (!b || !params || !params_valid(*params)) {
    set_error("mf_batch_set_params: invalid");
    return MF_ERR_INVALID;
  }
  if (params->reaction_policy != b->policy) {
    set_error("reaction_policy is fixed at construction");
    return MF_ERR_INVALID;
  }
  const bool mode_changed = params->mode != b->p.mode;
  const int old_bs = b->p.block_size;
  const int old_en = b->p.energy_diagnostics;
  const int old_fk = b->p.fp32_kernel;
  b->p = *params;
  if (mode_changed) {
    b->have_prev = false;
  }
  if (old_bs != params->block_size || old_en != params->energy_diagnostics ||
      old_fk != params->fp32_kernel)
    for (int prec = 0; prec < 2; ++prec) {
      int e = make_plan(b, prec);
      if (e != MF_OK) return e;
    }
  return build_identity_tables(b);
}

int mf_batch_set_graph(mf_batch *b, int32_t enable) {
  if (!b) return MF_ERR_INVALID;
  if (!enable) b->drop_graphs();
  b->use_graph = enable != 0;
  return MF_OK;
}

int mf_batch_autotune(mf_batch *b, int32_t precision, int32_t steps, const char *cache_path,
                      int32_t *n_timed) {
  if (!b || (precision != MF_PREC_FP64 && precision != MF_PREC_FP32C) || steps < 1 ||
      steps > 1000000) {
    set_error("mf_batch_autotune: bad arguments");
    return MF_ERR_INVALID;
  }
  if (b->p.block_size || b->p.fp32_kernel) {
    set_error("mf_batch_autotune: block_size or fp32_kernel forces the variant; set them to 0");
    return MF_ERR_INVALID;
  }
  const int prec = precision;
  int nv = 0;
  const Variant *vs = variants(&nv);
  std::map<TuneKey, std::vector<int>> buckets;
  std::vector<char> lin(b->R);
  for (int r = 0; r < b->R; ++r) {
    lin[r] = all_linear(b->host[r]);
    buckets[tune_key(prec, b->host[r], lin[r])].push_back(r);
  }
  // cache lines: prec lin regok lanes64 n256 block npt flags us_per_solve
  std::map<TuneKey, int> cached;
  if (cache_path) {
    std::ifstream f(cache_path);
    std::string line;
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream is(line);
      TuneKey k;
      int blk, npt, fl;
      double us;
      if (!(is >> k.prec >> k.lin >> k.regok >> k.lanes64 >> k.n256 >> blk >> npt >> fl >> us))
        continue;
      for (int i = 0; i < nv; ++i)
        if (vs[i].prec == k.prec && vs[i].block == blk && vs[i].npt == npt && vs[i].flags == fl)
          cached[k] = i;
    }
  }
  const mf_solver_params p0 = b->p;
  const std::vector<int> plan0 = b->plan[prec];
  const std::vector<size_t> lds0 = b->lds[prec];
  const std::map<TuneKey, int> tuned0 = b->tuned;
  b->p.mode = MF_MODE_RAMP;
  b->p.n_steps = steps;
  int timed = 0;
  std::ostringstream added;
  int err = MF_OK;
  for (const auto &kv : buckets) {
    const TuneKey &key = kv.first;
    const std::vector<int> &L = kv.second;
    auto fits_all = [&](int i) {
      for (int r : L)
        if (!variant_fits(b, prec, vs[i], b->host[r], lin[r])) return false;
      return true;
    };
    auto c = cached.find(key);
    if (c != cached.end() && fits_all(c->second)) {
      b->tuned[key] = c->second;
      continue;
    }
    // the global-state variant only where nothing else fits
    std::vector<int> cand;
    for (int i = 0; i < nv; ++i)
      if (!(vs[i].flags & kGlobalState) && fits_all(i)) cand.push_back(i);
    for (int i = 0; cand.empty() && i < nv; ++i)
      if ((vs[i].flags & kGlobalState) && fits_all(i)) cand.push_back(i);
    if (cand.empty()) continue;
    int best = cand[0];
    double best_us = 0.0;
    if (cand.size() > 1) {
      // enough solves to fill the GPU a few times over
      const int ns = std::max(512, std::min((int)L.size(), 2048));
      SolveSet S;
      S.prec = prec;
      S.rve.resize(ns);
      S.flags.assign(ns, 0);
      S.F.assign(9 * (size_t)ns, 0.0);
      for (int i = 0; i < ns; ++i) {
        S.rve[i] = L[i % L.size()];
        S.F[9 * i] = 1.01;
        S.F[9 * i + 4] = 1.0;
        S.F[9 * i + 8] = 1.0;
      }
      best = -1;
      for (int i : cand) {
        for (int r : L) {
          b->plan[prec][r] = i;
          b->lds[prec][r] = lds_bytes(vs[i], b->host[r].N, b->host[r].Nb, b->host[r].M,
                                      b->host[r].ell_free_entries);
        }
        const int cidx = b->cur;
        double us = 0.0;
        for (int rep = 0; rep < 2 && err == MF_OK; ++rep) {  // warm-up, then timed
          err = launch_set(b, S, b->xs[cidx].p, b->us[cidx].p, b->xs[cidx].p, b->us[cidx].p);
          if (err == MF_OK) err = finish_timing(b);
          us = 1e3 * b->last_ms / ns;
        }
        if (err != MF_OK) break;
        ++timed;
        if (best < 0 || us < best_us) {
          best = i;
          best_us = us;
        }
      }
      if (err != MF_OK) break;
    }
    b->tuned[key] = best;
    added << key.prec << ' ' << key.lin << ' ' << key.regok << ' ' << key.lanes64 << ' '
          << key.n256 << ' ' << vs[best].block << ' ' << vs[best].npt << ' ' << vs[best].flags
          << ' ' << best_us << '\n';
  }
  // Whole-batch check (also for buckets read from the cache): buckets timed
  // alone can mispredict the batch, whose launch groups run side by side;
  // keep the tuned buckets only if the full batch (one solve per RVE) runs
  // faster with them than with the plan they replace.  The check runs 4x
  // the candidate steps, and at least the batch's own ramp length up to 2000
  // steps: with 100-step candidates a 400-step check kept a plan that ran 1%
  // slower at 10000 steps.  The per-bucket timings go to the cache either
  // way.
  const int check_steps =
      std::max(4 * steps, p0.mode == MF_MODE_RAMP ? std::min(p0.n_steps, 2000) : 0);
  if (err == MF_OK && b->tuned != tuned0) {
    SolveSet S;
    S.prec = prec;
    S.rve.resize(b->R);
    S.flags.assign(b->R, 0);
    S.F.assign(9 * (size_t)b->R, 0.0);
    for (int r = 0; r < b->R; ++r) {
      S.rve[r] = r;
      S.F[9 * r] = 1.01;
      S.F[9 * r + 4] = 1.0;
      S.F[9 * r + 8] = 1.0;
    }
    b->p.n_steps = check_steps;
    auto batch_ms = [&](double &ms) {
      int e = make_plan(b, prec);
      const int cidx = b->cur;
      for (int rep = 0; rep < 2 && e == MF_OK; ++rep) {  // warm-up, then timed
        e = launch_set(b, S, b->xs[cidx].p, b->us[cidx].p, b->xs[cidx].p, b->us[cidx].p);
        if (e == MF_OK) e = finish_timing(b);
      }
      ms = b->last_ms;
      return e;
    };
    const std::map<TuneKey, int> tuned1 = b->tuned;
    double ms_old = 0.0, ms_new = 0.0;
    b->tuned = tuned0;
    err = batch_ms(ms_old);
    b->tuned = tuned1;
    if (err == MF_OK) err = batch_ms(ms_new);
    if (!(err == MF_OK && ms_new < ms_old)) b->tuned = tuned0;
  }
  b->p = p0;
  b->plan[prec] = plan0;
  b->lds[prec] = lds0;
  if (err != MF_OK) {
    b->tuned = tuned0;
    return err;
  }
  int e = make_plan(b, prec);
  if (e != MF_OK) return e;
  if ((e = build_identity_tables(b)) != MF_OK) return e;
  // Step chunks per solve (ramp FP32c runs, register-kernel groups): the
  // model's count and 1-4 forced are timed on the batch's own run path
  // (trial runs of the check's length, which leave the state untouched)
  // and the fastest is kept; params.chunks, when set, still wins.
  if (prec == MF_PREC_FP32C && p0.precision == MF_PREC_FP32C && p0.mode == MF_MODE_RAMP &&
      p0.chunks == 0) {
    bool chunkable = false;
    for (size_t gi = 0; gi < b->id_groups[prec].size(); ++gi) {
      const Variant &v = vs[b->id_groups[prec][gi].first];
      chunkable = chunkable || ((v.flags & kEllInRegs) && !(v.flags & kWaveRve));
    }
    if (chunkable) {
      const int tc0 = b->tuned_chunks;
      b->p.n_steps = check_steps;
      const double Fi[9] = {1.01, 0, 0, 0, 1, 0, 0, 0, 1};
      MF_HIP(b->F_in.ensure(9));
      MF_HIP(hipMemcpyAsync(b->F_in.p, Fi, sizeof(Fi), hipMemcpyHostToDevice, b->stream));
      int best_c = tc0;
      double best_ms = 0.0;
      for (int c = 0; c <= 4 && e == MF_OK; ++c) {
        b->tuned_chunks = c;
        if ((e = build_identity_tables(b)) != MF_OK) break;
        double ms = 0.0;
        for (int rep = 0; rep < 2 && e == MF_OK; ++rep) {  // warm-up, then timed
          e = run_device_fast(b, b->F_in.p, 1, 0, nullptr, nullptr);
          if (e == MF_OK) e = finish_timing(b);
          ms = b->last_ms;
        }
        if (e == MF_OK && (c == 0 || ms < best_ms)) {
          best_c = c;
          best_ms = ms;
        }
      }
      b->p = p0;
      b->tuned_chunks = e == MF_OK ? best_c : tc0;
      const int e2 = build_identity_tables(b);
      if (e != MF_OK) return e;
      if ((e = e2) != MF_OK) return e;
    }
  }
  if (cache_path && !added.str().empty()) {
    const bool fresh = !std::ifstream(cache_path).good();
    std::ofstream f(cache_path, std::ios::app);
    if (!f) {
      set_error(std::string("mf_batch_autotune: cannot write ") + cache_path);
      return MF_ERR_IO;
    }
    if (fresh)
      f << "# mumfim_amd autotune cache: prec lin regok lanes/64 nodes/256 block npt flags "
           "us_per_solve (gfx950)\n";
    f << added.str();
  }
  if (n_timed) *n_timed = timed;
  return MF_OK;
}

int mf_batch_last_energies(mf_batch *b, double *W) {
  if (!b || !W) {
    set_error("mf_batch_last_energies: NULL argument");
    return MF_ERR_INVALID;
  }
  if (!b->have_energy) {
    set_error("mf_batch_last_energies: the last run had no FP64 energy diagnostics");
    return MF_ERR_STATE;
  }
  MF_HIP(hipMemcpyAsync(W, b->energy.p, sizeof(double) * 4 * b->R, hipMemcpyDeviceToHost,
                        b->stream));
  MF_HIP(hipStreamSynchronize(b->stream));
  return MF_OK;
}

int mf_batch_set_damping(mf_batch *b, const double *eta) {
  if (!b) {
    set_error("mf_batch_set_damping: NULL batch");
    return MF_ERR_INVALID;
  }
  if (!eta) {
    b->drop_graphs();
    b->eta_rve.release();
    return MF_OK;
  }
  for (int r = 0; r < b->R; ++r)
    if (!(eta[r] >= 0.0) || !std::isfinite(eta[r])) {
      set_error("mf_batch_set_damping: damping must be finite and >= 0");
      return MF_ERR_INVALID;
    }
  std::vector<double> v(eta, eta + b->R);
  b->drop_graphs();
  MF_HIP(b->eta_rve.upload(v, b->stream));
  MF_HIP(hipStreamSynchronize(b->stream));
  return MF_OK;
}

int mf_batch_get_params(const mf_batch *b, mf_solver_params *params) {
  if (!b || !params) return MF_ERR_INVALID;
  *params = b->p;
  return MF_OK;
}

int mf_batch_run(mf_batch *b, const double *F, int32_t nF, double *sigma,
                 int32_t update_coords, int32_t *status) {
  if (!b || !F || !sigma) {
    set_error("mf_batch_run: NULL argument");
    return MF_ERR_INVALID;
  }
  int e = check_F(b, nF, update_coords);
  if (e != MF_OK) return e;
  MF_HIP(b->F_in.ensure(9 * (size_t)nF));
  MF_HIP(hipMemcpyAsync(b->F_in.p, F, sizeof(double) * 9 * nF,
                        hipMemcpyHostToDevice, b->stream));
  e = run_device_fast(b, b->F_in.p, nF, update_coords, nullptr, nullptr);
  if (e != MF_OK) return e;
  if (update_coords && b->p.mode == MF_MODE_RAMP) {
    b->lastF.assign(9 * (size_t)b->R, 0.0);
    for (int r = 0; r < b->R; ++r)
      std::copy(F + 9 * (nF == 1 ? 0 : r), F + 9 * (nF == 1 ? 0 : r) + 9,
                b->lastF.begin() + 9 * (size_t)r);
    b->lastF_on_device = false;
  }
  e = finish_timing(b);
  if (e != MF_OK) return e;
  return copy_out(b, sigma, status, b->R);
}

int mf_batch_run_device_async(mf_batch *b, const double *F_dev, int32_t nF,
                              double *sigma_dev, int32_t update_coords,
                              int32_t *status_dev) {
  if (!b || !F_dev || !sigma_dev) {
    set_error("mf_batch_run_device_async: NULL argument");
    return MF_ERR_INVALID;
  }
  int e = check_F(b, nF, update_coords);
  if (e != MF_OK) return e;
  if (!b->use_graph) return run_device_fast(b, F_dev, nF, update_coords, sigma_dev, status_dev);
  return run_device_graph(b, F_dev, nF, update_coords, sigma_dev, status_dev);
}

int mf_batch_run_device_noalloc(mf_batch *b, const double *F_dev, int32_t nF,
                                double *sigma_dev, int32_t update_coords,
                                int32_t *status_dev) {
  if (!b || !F_dev || !sigma_dev) {
    set_error("mf_batch_run_device_noalloc: NULL argument");
    return MF_ERR_INVALID;
  }
  int e = check_F(b, nF, update_coords);
  if (e != MF_OK) return e;
  const int R = b->R, prec = b->p.precision;
  if (prec == MF_PREC_FP64 && b->p.energy_diagnostics) {
    set_error("mf_batch_run_device_noalloc: energy diagnostics are not supported "
              "(they need a library buffer)");
    return MF_ERR_UNSUPPORTED;
  }
  // layout: F staging (9R doubles), step counters (R), statuses (R), global state
  const size_t off_steps = align256(sizeof(double) * 9 * (size_t)R);
  const size_t off_status = off_steps + align256(sizeof(int32_t) * (size_t)R);
  const size_t off_gstate = off_status + align256(sizeof(int32_t) * (size_t)R);
  const size_t need = off_gstate + sizeof(double) * b->id_gtotal[prec];
  if (need > (size_t)MF_NOALLOC_SCRATCH_BYTES) {
    set_error("mf_batch_run_device_noalloc: the batch needs " + std::to_string(need) +
              " bytes of scratch, above the maximum of " +
              std::to_string((size_t)MF_NOALLOC_SCRATCH_BYTES) +
              " (use mf_batch_run_device_async)");
    return MF_ERR_INVALID;
  }
  int prev = -1;
  MF_HIP(hipGetDevice(&prev));
  if (prev != b->device) MF_HIP(hipSetDevice(b->device));
  unsigned char *base = nullptr;
  e = scratch_address(b->device, &base);
  if (e == MF_OK) {
    RunScratch na{reinterpret_cast<double *>(base),
                  reinterpret_cast<int32_t *>(base + off_steps),
                  reinterpret_cast<int32_t *>(base + off_status),
                  reinterpret_cast<double *>(base + off_gstate)};
    e = run_device_fast(b, F_dev, nF, update_coords, sigma_dev, status_dev, &na);
  }
  if (prev != b->device) (void)hipSetDevice(prev);
  return e;
}

int mf_batch_synchronize(mf_batch *b) {
  if (!b) return MF_ERR_INVALID;
  MF_HIP(hipStreamSynchronize(b->stream));
  return finish_timing(b);
}

void *mf_batch_stream(mf_batch *b) { return b ? (void *)b->stream : nullptr; }

int mf_batch_last_solve_ms(mf_batch *b, double *ms, int32_t *n_launches) {
  if (!b) return MF_ERR_INVALID;
  if (ms) *ms = b->last_ms;
  if (n_launches) *n_launches = b->last_launches;
  return MF_OK;
}

int mf_batch_tangent(mf_batch *b, double *C, int32_t *status) {
  if (!b || !C) {
    set_error("mf_batch_tangent: NULL argument");
    return MF_ERR_INVALID;
  }
  DevBuf<double> Cd;
  MF_HIP(Cd.alloc(36 * (size_t)b->R));
  std::vector<int32_t> st;
  int e = tangent_impl(b, Cd.p, status ? &st : nullptr);
  if (e != MF_OK) return e;
  MF_HIP(hipMemcpyAsync(C, Cd.p, sizeof(double) * 36 * b->R,
                        hipMemcpyDeviceToHost, b->stream));
  MF_HIP(hipStreamSynchronize(b->stream));
  if (status) std::copy(st.begin(), st.end(), status);
  return MF_OK;
}

int mf_batch_run_with_tangent(mf_batch *b, const double *F, int32_t nF,
                              double *sigma, double *C, int32_t *status) {
  if (!b || !F || !sigma || !C) {
    set_error("mf_batch_run_with_tangent: NULL argument");
    return MF_ERR_INVALID;
  }
  int e = check_F(b, nF, 1);
  if (e != MF_OK) return e;
  const int R = b->R;
  hipStream_t s = b->stream;
  if (b->p.mode == MF_MODE_RELAX || b->p.precision != b->p.tangent_precision) {
    e = mf_batch_run(b, F, nF, sigma, 1, status);
    if (e != MF_OK) return e;
    std::vector<int32_t> st(R);
    e = mf_batch_tangent(b, C, status ? st.data() : nullptr);
    if (e != MF_OK) return e;
    if (status)
      for (int r = 0; r < R; ++r)
        if (status[r] == MF_STATUS_OK) status[r] = st[r];
    return MF_OK;
  }
  // ramp mode: base (committed) + perturbed solves in one launch
  MF_HIP(b->F_in.ensure(9 * (size_t)nF));
  MF_HIP(hipMemcpyAsync(b->F_in.p, F, sizeof(double) * 9 * nF,
                        hipMemcpyHostToDevice, s));
  hipError_t he = launch_update_volume(b->F_in.p, nF, b->V.p, R, s);
  if (he != hipSuccess) {
    set_error(hipGetErrorString(he));
    return MF_ERR_HIP;
  }
  const bool central = b->p.tangent_scheme == MF_TANGENT_CENTRAL;
  const bool tl = b->p.tangent_scheme == MF_TANGENT_LINEAR;
  const int per = tl ? 6 : central ? 13 : 7;
  const double h = b->p.fd_step;
  SolveSet S;
  S.prec = b->p.precision;
  S.tl = tl;
  for (int r = 0; r < R; ++r) {
    const double *F0 = F + 9 * (nF == 1 ? 0 : r);
    if (tl) {
      // solve 0 of each RVE commits the state; all six carry the stress
      for (int i = 0; i < 6; ++i) {
        double E[9], G[9];
        perturbation_rate(i, E);
        matmul3(E, F0, G);
        S.rve.push_back(r);
        S.flags.push_back(i == 0 ? (kWriteX | kWriteU) : 0);
        S.F.insert(S.F.end(), F0, F0 + 9);
        S.G.insert(S.G.end(), G, G + 9);
      }
      continue;
    }
    S.rve.push_back(r);
    S.flags.push_back(kWriteX | kWriteU);
    S.F.insert(S.F.end(), F0, F0 + 9);
    for (int k = 1; k < per; ++k) {
      int i = central ? (k - 1) / 2 : k - 1;
      double sgn = (central && ((k - 1) % 2 == 1)) ? -1.0 : 1.0;
      double Fp[9], Ft[9];
      perturbation(i, sgn * h, Fp);
      matmul3(Fp, F0, Ft);
      S.rve.push_back(r);
      S.flags.push_back(0);
      S.F.insert(S.F.end(), Ft, Ft + 9);
    }
  }
  const int c = b->cur;
  e = launch_set(b, S, b->xs[c].p, b->us[c].p, b->xs[c ^ 1].p, b->us[c ^ 1].p);
  if (e != MF_OK) return e;
  b->have_energy = false;  // this run's solves track no energies
  DevBuf<double> Cd;
  MF_HIP(Cd.alloc(36 * (size_t)R));
  he = tl ? launch_tl_assemble(b->dsigma.p, Cd.p, R, s)
          : launch_fd_all(b->sigma.p, Cd.p, R, central ? MF_TANGENT_CENTRAL : MF_TANGENT_FORWARD,
                          h, s);
  if (he != hipSuccess) {
    set_error(hipGetErrorString(he));
    return MF_ERR_HIP;
  }
  MF_HIP(hipMemcpy2DAsync(b->cur_stress.p, sizeof(double) * 6, b->sigma.p,
                          sizeof(double) * 6 * per, sizeof(double) * 6, R,
                          hipMemcpyDeviceToDevice, s));
  MF_HIP(hipMemcpy2DAsync(sigma, sizeof(double) * 6, b->sigma.p,
                          sizeof(double) * 6 * per, sizeof(double) * 6, R,
                          hipMemcpyDeviceToHost, s));
  MF_HIP(hipMemcpyAsync(C, Cd.p, sizeof(double) * 36 * R, hipMemcpyDeviceToHost, s));
  std::vector<int32_t> st(S.rve.size());
  MF_HIP(hipMemcpyAsync(st.data(), b->status.p, sizeof(int32_t) * st.size(),
                        hipMemcpyDeviceToHost, s));
  e = finish_timing(b);
  if (e != MF_OK) return e;
  MF_HIP(hipStreamSynchronize(s));
  b->have_run = true;
  b->have_prev = true;
  b->at_reference = false;
  b->cur ^= 1;
  b->lastF_on_device = false;
  b->lastF.assign(9 * (size_t)R, 0.0);
  for (int r = 0; r < R; ++r)
    std::copy(F + 9 * (nF == 1 ? 0 : r), F + 9 * (nF == 1 ? 0 : r) + 9,
              b->lastF.begin() + 9 * (size_t)r);
  if (status) {
    for (int r = 0; r < R; ++r) status[r] = MF_STATUS_OK;
    for (size_t k = 0; k < st.size(); ++k)
      if (status[S.rve[k]] == MF_STATUS_OK) status[S.rve[k]] = st[k];
  }
  return MF_OK;
}

int mf_batch_orientation_3d(mf_batch *b, double *omega) {
  return mf_batch_orientation_2d(b, nullptr, 0, omega);
}

int mf_batch_orientation_2d(mf_batch *b, const double *normal, int32_t n_normal,
                            double *omega) {
  if (!b || !omega) return MF_ERR_INVALID;
  if (normal && n_normal != 1 && n_normal != b->R) {
    set_error("normal must hold 1 or n_rve vectors");
    return MF_ERR_INVALID;
  }
  hipStream_t s = b->stream;
  MF_HIP(b->omega.ensure(9 * (size_t)b->R));
  const double *nd = nullptr;
  if (normal) {
    MF_HIP(b->normal.ensure(3 * (size_t)n_normal));
    MF_HIP(hipMemcpyAsync(b->normal.p, normal, sizeof(double) * 3 * n_normal,
                          hipMemcpyHostToDevice, s));
    nd = b->normal.p;
  }
  hipError_t e = launch_orientation(b->rves.p, b->xs[b->cur].p, b->fib.p, nd,
                                    n_normal, b->omega.p, b->R, s);
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return MF_ERR_HIP;
  }
  MF_HIP(hipMemcpyAsync(omega, b->omega.p, sizeof(double) * 9 * b->R,
                        hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  return MF_OK;
}

int64_t mf_batch_total_nodes(const mf_batch *b) { return b ? b->total_nodes : 0; }

int mf_batch_get_state(mf_batch *b, double *x, double *u, double *volume) {
  if (!b) return MF_ERR_INVALID;
  const size_t nn = 3 * (size_t)b->total_nodes;
  std::vector<double> xd(nn), ud(nn);
  hipStream_t s = b->stream;
  MF_HIP(hipMemcpyAsync(xd.data(), b->xs[b->cur].p, sizeof(double) * nn,
                        hipMemcpyDeviceToHost, s));
  MF_HIP(hipMemcpyAsync(ud.data(), b->us[b->cur].p, sizeof(double) * nn,
                        hipMemcpyDeviceToHost, s));
  if (volume)
    MF_HIP(hipMemcpyAsync(volume, b->V.p, sizeof(double) * b->R,
                          hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < b->R; ++r) {
    const PackedRVE &p = b->host[r];
    const int64_t off = b->rdev[r].node_off;
    for (int i = 0; i < p.N; ++i) {
      const int64_t src = off + p.rank_of_input[i], dst = off + i;
      for (int d = 0; d < 3; ++d) {
        if (x) x[3 * dst + d] = xd[3 * src + d];
        if (u) u[3 * dst + d] = ud[3 * src + d];
      }
    }
  }
  return MF_OK;
}

int mf_batch_set_state(mf_batch *b, const double *x, const double *u,
                       const double *volume) {
  if (!b) return MF_ERR_INVALID;
  const size_t nn = 3 * (size_t)b->total_nodes;
  std::vector<double> xd(nn), ud(nn);
  for (int r = 0; r < b->R; ++r) {
    const PackedRVE &p = b->host[r];
    const int64_t off = b->rdev[r].node_off;
    for (int i = 0; i < p.N; ++i) {
      const int64_t dst = off + p.rank_of_input[i], src = off + i;
      for (int d = 0; d < 3; ++d) {
        xd[3 * dst + d] = x ? x[3 * src + d] : 0.0;
        ud[3 * dst + d] = u ? u[3 * src + d] : 0.0;
      }
    }
  }
  hipStream_t s = b->stream;
  if (x)
    MF_HIP(hipMemcpyAsync(b->xs[b->cur].p, xd.data(), sizeof(double) * nn,
                          hipMemcpyHostToDevice, s));
  if (u)
    MF_HIP(hipMemcpyAsync(b->us[b->cur].p, ud.data(), sizeof(double) * nn,
                          hipMemcpyHostToDevice, s));
  if (volume)
    MF_HIP(hipMemcpyAsync(b->V.p, volume, sizeof(double) * b->R,
                          hipMemcpyHostToDevice, s));
  MF_HIP(hipStreamSynchronize(s));
  b->have_prev = false;
  b->at_reference = false;
  return MF_OK;
}

int mf_batch_rve_sizes(const mf_batch *b, int32_t *sizes) {
  if (!b || !sizes) return MF_ERR_INVALID;
  for (int r = 0; r < b->R; ++r) {
    const PackedRVE &p = b->host[r];
    sizes[4 * r + 0] = p.N;
    sizes[4 * r + 1] = p.Nf;
    sizes[4 * r + 2] = p.Nb;
    sizes[4 * r + 3] = p.M;
  }
  return MF_OK;
}

int mf_batch_plan(mf_batch *b, int32_t *plan) {
  if (!b) return MF_ERR_INVALID;
  int nv = 0;
  const Variant *vs = variants(&nv);
  const int prec = b->p.precision;
  // step chunks of each RVE's launch group (one-solve-per-RVE runs)
  std::vector<int> nch(b->R, 1);
  {
    std::vector<int32_t> order(b->R);
    if (b->R > 0)
      MF_HIP(hipMemcpy(order.data(), b->id_order[prec].p, sizeof(int32_t) * b->R,
                       hipMemcpyDeviceToHost));
    int off = 0;
    for (size_t gi = 0; gi < b->id_groups[prec].size(); ++gi) {
      for (int k = 0; k < b->id_groups[prec][gi].second; ++k)
        nch[order[off + k]] = b->id_group_chunks[prec][gi];
      off += b->id_groups[prec][gi].second;
    }
  }
  std::vector<int> seen;
  for (int r = 0; r < b->R; ++r) {
    const Variant &v = vs[b->plan[prec][r]];
    if (plan) {
      plan[4 * r + 0] = v.block;
      plan[4 * r + 1] = v.npt;
      plan[4 * r + 2] = v.flags | (nch[r] > 1 ? nch[r] << 24 : 0);
      plan[4 * r + 3] = (int32_t)b->lds[prec][r];
    }
    if (std::find(seen.begin(), seen.end(), b->plan[prec][r]) == seen.end())
      seen.push_back(b->plan[prec][r]);
  }
  return (int)seen.size();
}

int mf_generate_network(const mf_gen_params *g, int32_t *n_nodes,
                        int32_t *n_fibers, int32_t *n_boundary, double *coords,
                        int32_t *fibers, int32_t *fiber_reaction,
                        int32_t *boundary, mf_reaction *reactions) {
  if (!g || !n_nodes || !n_fibers || !n_boundary) {
    set_error("mf_generate_network: NULL argument");
    return MF_ERR_INVALID;
  }
  Network net;
  int e = generate_network(*g, net);
  if (e != MF_OK) return e;
  *n_nodes = net.N;
  *n_fibers = net.M;
  *n_boundary = net.Nb;
  if (coords) std::copy(net.X.begin(), net.X.end(), coords);
  if (fibers) std::copy(net.fib.begin(), net.fib.end(), fibers);
  if (fiber_reaction) std::copy(net.rid.begin(), net.rid.end(), fiber_reaction);
  if (boundary) std::copy(net.bnd.begin(), net.bnd.end(), boundary);
  if (reactions)
    for (size_t i = 0; i < net.reactions.size(); ++i) {
      const Reaction &q = net.reactions[i];
      mf_reaction &r = reactions[i];
      std::memset(&r, 0, sizeof(r));
      r.type = q.type;
      r.radius = q.radius;
      r.E = q.E;
      r.density = q.rho;
      r.B = q.B;
      r.length_ratio_trns = q.lt;
      r.cmp_ratio_trns = q.lc;
    }
  return MF_OK;
}

int mf_device_alloc(void **ptr, int64_t bytes) {
  if (!ptr || bytes < 0) return MF_ERR_INVALID;
  *ptr = nullptr;
  if (bytes == 0) return MF_OK;
  MF_HIP(hipMalloc(ptr, (size_t)bytes));
  return MF_OK;
}

int mf_device_free(void *ptr) {
  if (ptr) MF_HIP(hipFree(ptr));
  return MF_OK;
}

int mf_memcpy_to_device(void *dst, const void *src, int64_t bytes) {
  if ((!dst || !src) && bytes) return MF_ERR_INVALID;
  if (bytes) MF_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return MF_OK;
}

int mf_memcpy_to_host(void *dst, const void *src, int64_t bytes) {
  if ((!dst || !src) && bytes) return MF_ERR_INVALID;
  if (bytes) MF_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return MF_OK;
}

int mf_device_synchronize(void) {
  MF_HIP(hipDeviceSynchronize());
  return MF_OK;
}

int mf_measure_lds_bandwidth(double *bytes_per_s) {
  if (!bytes_per_s) return MF_ERR_INVALID;
  MF_HIP(lds_probe(bytes_per_s));
  return MF_OK;
}

int mf_measure_hbm_bandwidth(int64_t bytes, double *bytes_per_s) {
  if (!bytes_per_s || bytes < (1 << 20)) return MF_ERR_INVALID;
  MF_HIP(hbm_probe(bytes, bytes_per_s));
  return MF_OK;
}

int mf_uniform(uint64_t seed, uint64_t stream, int64_t n, double *out) {
  if (!out || n < 0) return MF_ERR_INVALID;
  SplitMix64 rng(seed, stream);
  for (int64_t i = 0; i < n; ++i) out[i] = rng.uniform();
  return MF_OK;
}

int mf_load_network(const char *path, int32_t *n_nodes, int32_t *n_fibers,
                    int32_t *n_boundary, int32_t *n_reactions,
                    int32_t max_reactions, double *coords, int32_t *fibers,
                    int32_t *fiber_reaction, int32_t *boundary,
                    mf_reaction *reactions) {
  if (!path || !n_nodes || !n_fibers || !n_boundary || !n_reactions) {
    set_error("mf_load_network: NULL argument");
    return MF_ERR_INVALID;
  }
  Network net;
  int e = load_network_file(path, net);
  if (e != MF_OK) return e;
  *n_nodes = net.N;
  *n_fibers = net.M;
  *n_boundary = net.Nb;
  *n_reactions = (int32_t)net.reactions.size();
  if (coords) std::copy(net.X.begin(), net.X.end(), coords);
  if (fibers) std::copy(net.fib.begin(), net.fib.end(), fibers);
  if (fiber_reaction) std::copy(net.rid.begin(), net.rid.end(), fiber_reaction);
  if (boundary) std::copy(net.bnd.begin(), net.bnd.end(), boundary);
  if (reactions) {
    if ((int)net.reactions.size() > max_reactions) {
      set_error("reaction buffer too small");
      return MF_ERR_INVALID;
    }
    for (size_t i = 0; i < net.reactions.size(); ++i) {
      const Reaction &q = net.reactions[i];
      mf_reaction &r = reactions[i];
      std::memset(&r, 0, sizeof(r));
      r.type = q.type;
      r.radius = q.radius;
      r.E = q.E;
      r.density = q.rho;
      r.B = q.B;
      r.length_ratio_trns = q.lt;
      r.cmp_ratio_trns = q.lc;
    }
  }
  return MF_OK;
}


// ---- surrogate path --------------------------------------------------------

static int surrogate_prepare(mf_batch *b, const double *F, int32_t nF) {
  const int R = b->R;
  hipStream_t s = b->stream;
  if (!b->sur_desc.p) {
    // node count, fibers per volume (V = 1 at the reference state), mean
    // fiber length, nonlinear-fiber fraction
    std::vector<float> d(4 * (size_t)R);
    for (int r = 0; r < R; ++r) {
      const PackedRVE &p = b->host[r];
      d[4 * r + 0] = (float)p.N;
      d[4 * r + 1] = (float)p.M;
      d[4 * r + 2] = (float)p.mean_l0;
      d[4 * r + 3] = (float)p.nonlinear_frac;
    }
    MF_HIP(b->sur_desc.upload(d, s));
  }
  MF_HIP(b->F_in.ensure(9 * (size_t)nF));
  MF_HIP(hipMemcpyAsync(b->F_in.p, F, sizeof(double) * 9 * nF, hipMemcpyHostToDevice, s));
  MF_HIP(b->sur_X.ensure((size_t)R * MF_MLP_IN));
  MF_HIP(launch_surrogate_inputs(b->F_in.p, nF, b->sur_desc.p, R, b->sur_X.p, s));
  return MF_OK;
}

int mf_surrogate_inputs(mf_batch *b, const double *F, int32_t nF, float *X) {
  if (!b || !F || !X) {
    set_error("mf_surrogate_inputs: NULL argument");
    return MF_ERR_INVALID;
  }
  int e = check_F(b, nF, 0);
  if (e != MF_OK) return e;
  if ((e = surrogate_prepare(b, F, nF)) != MF_OK) return e;
  MF_HIP(hipMemcpyAsync(X, b->sur_X.p, sizeof(float) * MF_MLP_IN * b->R,
                        hipMemcpyDeviceToHost, b->stream));
  MF_HIP(hipStreamSynchronize(b->stream));
  return MF_OK;
}

int mf_surrogate_run(mf_batch *b, mf_mlp *m, const double *F, int32_t nF, double *sigma,
                     double *C, int32_t *status, mf_surrogate_stats *stats) {
  const auto t_start = std::chrono::steady_clock::now();
  if (!b || !m || !F || !sigma) {
    set_error("mf_surrogate_run: NULL argument");
    return MF_ERR_INVALID;
  }
  if (b->p.mode != MF_MODE_RAMP || !b->at_reference) {
    set_error("mf_surrogate_run: needs a ramp-mode batch in its reference configuration "
              "(the surrogate maps total strain from the reference state)");
    return MF_ERR_STATE;
  }
  int e = check_F(b, nF, 0);
  if (e != MF_OK) return e;
  const int R = b->R;
  hipStream_t s = b->stream;
  const int K = mlp_members(m);
  MF_HIP(hipEventRecord(b->ev0, s));
  if ((e = surrogate_prepare(b, F, nF)) != MF_OK) return e;
  MF_HIP(b->sur_Y.ensure((size_t)R * MF_MLP_OUT));
  MF_HIP(b->sur_spread.ensure((size_t)R * MF_MLP_OUT));
  MF_HIP(b->sur_flag.ensure(R));
  MF_HIP(b->sur_sigma.ensure(6 * (size_t)R));
  if (C) {
    MF_HIP(b->sur_J.ensure(36 * (size_t)R));
    MF_HIP(b->sur_C.ensure(36 * (size_t)R));
  }
  float *spread = K > 1 ? b->sur_spread.p : nullptr;
  if ((e = mlp_predict(m, -1, b->sur_X.p, R, b->sur_Y.p, C ? b->sur_J.p : nullptr, spread, s)) !=
      MF_OK)
    return e;
  MF_HIP(launch_surrogate_flags(b->sur_X.p, b->sur_Y.p, spread, R, mlp_estimator(m),
                                b->sur_flag.p, s));
  MF_HIP(launch_surrogate_outputs(b->F_in.p, nF, b->sur_Y.p, C ? b->sur_J.p : nullptr, R,
                                  b->sur_sigma.p, C ? b->sur_C.p : nullptr, s));
  MF_HIP(hipEventRecord(b->ev1, s));
  std::vector<int32_t> flag(R);
  MF_HIP(hipMemcpyAsync(flag.data(), b->sur_flag.p, sizeof(int32_t) * R, hipMemcpyDeviceToHost,
                        s));
  MF_HIP(hipMemcpyAsync(sigma, b->sur_sigma.p, sizeof(double) * 6 * R, hipMemcpyDeviceToHost,
                        s));
  if (C)
    MF_HIP(hipMemcpyAsync(C, b->sur_C.p, sizeof(double) * 36 * R, hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  float mlp_ms = 0;
  MF_HIP(hipEventElapsedTime(&mlp_ms, b->ev0, b->ev1));
  std::vector<int32_t> rej;
  int64_t n_oor = 0;
  for (int r = 0; r < R; ++r) {
    if (status) status[r] = MF_STATUS_OK;
    if (flag[r]) rej.push_back(r);
    if (flag[r] == 2) ++n_oor;
  }
  float fb_ms = 0;
  if (!rej.empty()) {
    // explicit trial solves of the rejected RVEs from the reference state;
    // with C: base + (F_i(+h) F, F_i(-h) F) for i = 0..5, 13 per RVE
    const int per = C ? 13 : 1;
    const int nr = (int)rej.size();
    SolveSet S;
    S.prec = C ? b->p.tangent_precision : b->p.precision;
    S.rve.resize((size_t)nr * per);
    S.flags.assign((size_t)nr * per, 0);
    S.F.resize(9 * (size_t)nr * per);
    const double h = b->p.fd_step;
    for (int k = 0; k < nr; ++k) {
      const int r = rej[k];
      const double *Fr = F + 9 * (nF == 1 ? 0 : r);
      for (int q = 0; q < per; ++q) {
        S.rve[(size_t)k * per + q] = r;
        double *dst = S.F.data() + 9 * ((size_t)k * per + q);
        if (q == 0) {
          std::copy(Fr, Fr + 9, dst);
        } else {
          double Fp[9];
          perturbation((q - 1) / 2, (q - 1) % 2 ? -h : h, Fp);
          matmul3(Fp, Fr, dst);
        }
      }
    }
    DevBuf<double> sig;
    DevBuf<int32_t> st;
    MF_HIP(sig.alloc(6 * S.rve.size()));
    MF_HIP(st.alloc(S.rve.size()));
    MF_HIP(hipEventRecord(b->ev0, s));
    const int c = b->cur;
    if ((e = launch_set(b, S, b->xs[c].p, b->us[c].p, b->xs[c].p, b->us[c].p, sig.p, st.p)) !=
        MF_OK)
      return e;
    DevBuf<double> Cd;
    if (C) {
      MF_HIP(Cd.alloc(36 * (size_t)nr));
      MF_HIP(launch_fd_all(sig.p, Cd.p, nr, MF_TANGENT_CENTRAL, h, s));
    }
    MF_HIP(hipEventRecord(b->ev1, s));
    std::vector<double> sh(6 * S.rve.size()), ch(C ? 36 * (size_t)nr : 0);
    std::vector<int32_t> sth(S.rve.size());
    MF_HIP(hipMemcpyAsync(sh.data(), sig.p, sizeof(double) * sh.size(), hipMemcpyDeviceToHost,
                          s));
    MF_HIP(hipMemcpyAsync(sth.data(), st.p, sizeof(int32_t) * sth.size(),
                          hipMemcpyDeviceToHost, s));
    if (C)
      MF_HIP(hipMemcpyAsync(ch.data(), Cd.p, sizeof(double) * ch.size(), hipMemcpyDeviceToHost,
                            s));
    MF_HIP(hipStreamSynchronize(s));
    MF_HIP(hipEventElapsedTime(&fb_ms, b->ev0, b->ev1));
    for (int k = 0; k < nr; ++k) {
      const int r = rej[k];
      std::copy(sh.begin() + 6 * (size_t)k * per, sh.begin() + 6 * (size_t)k * per + 6,
                sigma + 6 * (size_t)r);
      if (C) std::copy(ch.begin() + 36 * (size_t)k, ch.begin() + 36 * (size_t)k + 36,
                       C + 36 * (size_t)r);
      if (status) {
        int sv = MF_STATUS_FALLBACK;
        for (int q = 0; q < per; ++q)
          if (sth[(size_t)k * per + q] != MF_STATUS_OK) sv = sth[(size_t)k * per + q];
        status[r] = sv;
      }
    }
  }
  if (stats) {
    stats->n = R;
    stats->n_fallback = (int64_t)rej.size();
    stats->n_out_of_range = n_oor;
    stats->mlp_ms = mlp_ms;
    stats->fallback_ms = fb_ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t_start)
                          .count();
  }
  return MF_OK;
}

}  // extern "C"
