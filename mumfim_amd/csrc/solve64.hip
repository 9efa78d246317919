// FP64 explicit solve with the reference's operation order.  This file is
// compiled with -ffp-contract=off so every product and sum rounds as in the
// reference C++ (BatchedFiberRVEAnalysisExplicitTeamOuterLoop.h:180-437,
// BatchedFiberRVEAnalysisExplicitBase.h:155-287).
#include "kernels_common.hpp"

namespace mf {
namespace kern {

// ============================================================================
// FP64 kernel
// ============================================================================
// EN adds the energy accounting of the non-batched explicit solver
// (computeEnergies_ / computeKineticEnergy_ / checkEnergyBalance_,
// FiberRVEAnalysisExplicit_impl.h:1101-1138, 590-600): per step
// W_int += 1/2 du.(f_int_prev + f_int) over every node, W_ext the same over
// the prescribed boundary nodes (their reaction is their internal force),
// W_damp += 1/2 du.(f_damp_prev + f_damp) with f_damp = eta m v, and
// W_kin = 1/2 m v.v.
template <int B, int NPT, int FL, bool EN>
__global__ __launch_bounds__(B) void solve64(SolveArgs a) {
#pragma clang fp contract(off)
  constexpr bool kEll = (FL & kEllInLds) != 0;
  constexpr bool kDbuf = (FL & kDoubleBuf) != 0;
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const double eta_r = a.eta_rve ? a.eta_rve[rv] : a.eta;
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int64_t nb = R.node_off;

  // LDS carve (mirrors lds_bytes)
  size_t off = 0;
  double *pos = reinterpret_cast<double *>(smem);
  off += sizeof(double) * 3 * Np * (kDbuf ? 2 : 1);
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * 6 * Nbp;
  double *fbl = reinterpret_cast<double *>(smem + off);  // boundary f_int, last step
  off += EN ? sizeof(double) * 3 * Nbp : 0;
  ReactDev *rxs = reinterpret_cast<ReactDev *>(smem + off);
  off += align16(sizeof(ReactDev) * kMaxReactions);
  double *l0s = reinterpret_cast<double *>(smem + off);
  off += kEll ? sizeof(double) * R.ell_free_entries : 0;
  int32_t *ws = reinterpret_cast<int32_t *>(smem + off);

  const int32_t *gw = a.ell_w + R.ell_off;
  const double *gl0 = a.ell_l0 + R.ell_off;
  if (kEll) {
    for (int i = tid; i < R.ell_free_entries; i += B) {
      ws[i] = gw[i];
      l0s[i] = gl0[i];
    }
  }
  const int32_t *Ew = kEll ? ws : gw;
  const double *El0 = kEll ? l0s : gl0;
  if (tid < R.n_react) rxs[tid] = a.reacts[R.react_off + tid];
  const bool lin_only =
      R.n_react == 1 && a.reacts[R.react_off].type == MF_REACTION_LINEAR;
  const Lin64 lin{a.reacts[R.react_off].E, a.reacts[R.react_off].A};

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;

  // ---- initial state -------------------------------------------------------
  // relax: u += (F - I) x (applyIncrementalDeformationToDisplacement,
  //        ...Base.h:44-54); ramp: boundary increment g = (F - I) x kept aside.
  // x = X + u (getCurrentCoord, ...Base.h:192-202).
  double xr[NPT][3], vr[NPT][3], fr[NPT][3], mr[NPT];
  int degr[NPT], Dr[NPT], baser[NPT];
  auto init_node = [&](int r, double (&x)[3]) {
    const int64_t g = nb + r;
    double xp[3], up[3], u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = a.x_in[3 * g + d];
      up[d] = a.u_in[3 * g + d];
    }
    double inc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      inc[k] = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
               (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
               (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      u[k] = ramp ? up[k] : inc[k] + up[k];
      x[k] = a.X[3 * g + k] + u[k];
      pos[k * Np + r] = x[k];
      if (kDbuf) pos[3 * Np + k * Np + r] = x[k];
    }
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        bst[k * Nbp + b] = x[k];
        bst[(3 + k) * Nbp + b] = ramp ? inc[k] : 0.0;
      }
    }
  };
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    double x[3] = {0, 0, 0};
    if (r < N) init_node(r, x);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xr[s][d] = x[d];
      vr[s][d] = 0.0;
      fr[s][d] = 0.0;
    }
    degr[s] = 0;
    Dr[s] = 0;
    baser[s] = 0;
    mr[s] = 1.0;
    if (r < Nf) {
      mr[s] = a.mass[nb + r];
      degr[s] = a.deg[nb + r];
      const int k = r / kWave;
      Dr[s] = a.slice_D[R.slice_off + k];
      baser[s] = a.slice_base[R.slice_off + k] + lane;
    }
  }
  for (int r = NPT * B + tid; r < N; r += B) {
    double x[3];
    init_node(r, x);
  }
  __syncthreads();

  const double c_sound = R.c_sound;
  int cur = 0;
  // internal force of a free node at the positions in buffer `buf`
  auto node_force = [&](int s, int buf, double (&f)[3]) {
    const double *P = pos + (kDbuf ? buf * 3 * Np : 0);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int D = __builtin_amdgcn_readfirstlane(Dr[s]);
    // rolling prefetch: entry j+1's list word and l0 load while entry j is
    // evaluated (padding entries exist up to the slice width)
    int32_t wv_n = D > 0 ? Ew[baser[s]] : 0;
    double l0_n = D > 0 ? El0[baser[s]] : 1.0;
    for (int j = 0; j < D; ++j) {
      const int32_t wv = wv_n;
      const double l0 = l0_n;
      if (j + 1 < D) {
        wv_n = Ew[baser[s] + (j + 1) * kWave];
        l0_n = El0[baser[s] + (j + 1) * kWave];
      }
      if (j < degr[s]) {
        const int w = wv & 0xFFFFFF;
        const double dx = P[w] - xr[s][0];
        const double dy = P[Np + w] - xr[s][1];
        const double dz = P[2 * Np + w] - xr[s][2];
        const double l = sqrt(dx * dx + dy * dy + dz * dz);
        double frc;
        if (lin_only) {
          double lr = l / l0;
          double gs = 1.0 / 2.0 * (lr * lr - 1);
          frc = lr * lin.E * lin.A * gs;
        } else {
          frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
        }
        const double t = frc / l;
        fx = fx - dx * t;
        fy = fy - dy * t;
        fz = fz - dz * t;
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  // min fiber length over all fibers (getElementLengths, TeamOuterLoop.h:53-82)
  auto min_length = [&](int buf) {
    const double *P = pos + (kDbuf ? buf * 3 * Np : 0);
    double lmin = 1.0e300;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += B) {
      const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
      double x1 = P[n2] - P[n1];
      double x2 = P[Np + n2] - P[Np + n1];
      double x3 = P[2 * Np + n2] - P[2 * Np + n1];
      double l = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
      lmin = l < lmin ? l : lmin;
    }
    return lmin;
  };

  // internal force of boundary node Nf + b at the positions in buffer `buf`
  auto bnd_force = [&](int b, int buf, double (&f)[3]) {
    const double *P = pos + (kDbuf ? buf * 3 * Np : 0);
    const int r = Nf + b;
    const double xb0 = P[r], xb1 = P[Np + r], xb2 = P[2 * Np + r];
    const int base = a.slice_base[R.slice_off + r / kWave] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const double l0 = gl0[idx];
      const int w = wv & 0xFFFFFF;
      const double dx = P[w] - xb0;
      const double dy = P[Np + w] - xb1;
      const double dz = P[2 * Np + w] - xb2;
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      double frc;
      if (lin_only) {
        double lr = l / l0;
        double gs = 1.0 / 2.0 * (lr * lr - 1);
        frc = lr * lin.E * lin.A * gs;
      } else {
        frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
      }
      const double t = frc / l;
      f1 = f1 - dx * t;
      f2 = f2 - dy * t;
      f3 = f3 - dz * t;
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
  };
  // energy accumulators (sums of du.(f_prev + f), halved when reported)
  double w_int = 0.0, w_ext = 0.0, w_damp = 0.0;
  double fl[NPT][3], dl[NPT][3], du[NPT][3];

  // ---- RunInitialization (TeamOuterLoop.h:215-289) ---------------------------
  double fint[NPT][3];
  double resid_local = 0.0;
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    if (s * B + tid < Nf) {
      node_force(s, cur, fint[s]);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double mass = mr[s] * eta_r;
        double fdamp = mass * vr[s][d];
        fr[s][d] = -fint[s][d] - fdamp;
      }
    }
  }
  double lmin = min_length(cur);
  block_min_sum<B>(lmin, resid_local, red);
  double dt_crit = lmin / c_sound;
  if (EN) {
#pragma unroll
    for (int s = 0; s < NPT; ++s)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        fl[s][d] = s * B + tid < Nf ? fint[s][d] : 0.0;
        dl[s][d] = 0.0;
        du[s][d] = 0.0;
      }
    for (int b = tid; b < Nb; b += B) {
      double f[3];
      bnd_force(b, cur, f);
#pragma unroll
      for (int d = 0; d < 3; ++d) fbl[d * Nbp + b] = f[d];
    }
  }
  // totals of the four energies; a failed balance check stops the solve
  auto energy_totals = [&](double (&W)[4]) {
    double wk = 0.0;
#pragma unroll
    for (int s = 0; s < NPT; ++s)
      if (s * B + tid < Nf)
#pragma unroll
        for (int d = 0; d < 3; ++d) wk += mr[s] * vr[s][d] * vr[s][d];
    W[0] = w_int;
    W[1] = w_ext;
    W[2] = w_damp;
    W[3] = wk;
    __syncthreads();  // every wave is done with the previous use of red
    block_sum<B, 4>(W, red);
#pragma unroll
    for (int i = 0; i < 4; ++i) W[i] = 0.5 * W[i];
  };
  auto balance_ok = [&](const double (&W)[4]) {
    const double bal = W[3] + W[0] + W[2] - W[1];
    const double m = fmax(fmax(fabs(W[1]), fmax(fabs(W[0]), fabs(W[3]))), fabs(W[2]));
    return fabs(bal) <= a.energy_eps * m;
  };

  // ---- main loop (TeamOuterLoop.h:340-371) -----------------------------------
  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  for (;;) {
    const double h = a.crit * dt_crit * 0.5;
    const int nxt = kDbuf ? cur ^ 1 : 0;
    if (!kDbuf) __syncthreads();  // all force reads of the last step done
    double *Pn = pos + (kDbuf ? nxt * 3 * Np : 0);
    // updates(): half kick + drift of free nodes (TeamOuterLoop.h:157-175)
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double vl = vr[s][d] + h * fr[s][d] / mr[s];
          vr[s][d] = vl;
          xr[s][d] = xr[s][d] + 2 * h * vl;
          if (EN) du[s][d] = 2 * h * vl;
          Pn[d * Np + r] = xr[s][d];
        }
      }
    }
    if (ramp && tid < Nb) {  // waves without boundary nodes skip the amplitude
      const double alpha = ramp_alpha(step + 1, a.n_steps, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          Pn[d * Np + Nf + b] = bst[d * Nbp + b] + alpha * bst[(3 + d) * Nbp + b];
      }
    }
    __syncthreads();
    cur = nxt;
    const bool refresh = (step % freq) == 0;
    double rsum = 0.0;
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
        node_force(s, cur, fint[s]);
        double rn = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double mass = mr[s] * eta_r;
          double fdamp = mass * vr[s][d];
          double lres = -fint[s][d];
          fr[s][d] = lres - fdamp;
          rn += lres * lres;
          if (EN) {
            w_int += du[s][d] * (fl[s][d] + fint[s][d]);
            w_damp += du[s][d] * (dl[s][d] + fdamp);
            fl[s][d] = fint[s][d];
            dl[s][d] = fdamp;
          }
        }
        rsum += rn;
      }
    }
    if (EN && ramp) {
      // prescribed boundary: du_b from the amplitude, reaction = f_int
      const double a0 = step == 0 ? 0.0 : ramp_alpha(step, a.n_steps, a.amplitude);
      const double a1 = ramp_alpha(step + 1, a.n_steps, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
        double f[3];
        bnd_force(b, cur, f);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double x1 = bst[d * Nbp + b] + a1 * bst[(3 + d) * Nbp + b];
          const double x0 = bst[d * Nbp + b] + a0 * bst[(3 + d) * Nbp + b];
          const double t = (x1 - x0) * (fbl[d * Nbp + b] + f[d]);
          w_int += t;
          w_ext += t;
          fbl[d * Nbp + b] = f[d];
        }
      }
    }
    if (refresh) {
      double lm = min_length(cur);
      block_min_sum<B>(lm, rsum, red);
      dt_crit = lm / c_sound;
      residual = sqrt(rsum);
    }
    // updateAccelVels(): second half kick (TeamOuterLoop.h:144-156)
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d) vr[s][d] += h * fr[s][d] / mr[s];
      }
    }
    ++step;
    if (EN && a.energy_eps > 0.0 && step > 10 && step % freq == 0) {
      double W[4];
      energy_totals(W);
      if (!balance_ok(W)) {
        status = MF_STATUS_ENERGY;
        break;
      }
    }
    if (ramp) {
      if (step >= a.n_steps) break;
    } else {
      if (!(residual > a.tol)) break;
      if (step >= a.max_steps) {
        status = MF_STATUS_STEP_CAP;
        break;
      }
    }
  }
  if (residual != residual) status = MF_STATUS_NAN;
  if (EN) {
    double W[4];
    energy_totals(W);
    if (status == MF_STATUS_OK && a.energy_eps > 0.0 && step > 10 && !balance_ok(W))
      status = MF_STATUS_ENERGY;
    if (a.energy && tid == 0)
#pragma unroll
      for (int i = 0; i < 4; ++i) a.energy[4 * (int64_t)sid + i] = W[i];
  }

  // ---- FinalizeData + computeCauchyStress ------------------------------------
  // Boundary internal forces gather every incident fiber (TeamOuterLoop.h:
  // 374-434); stress = s/V * sum_b sym(x_b (x) f_b) (...Base.h:84-123).
  const double *P = pos + (kDbuf ? cur * 3 * Np : 0);
  __syncthreads();  // the last refresh reduction has been read by every wave
  double st[6] = {0, 0, 0, 0, 0, 0};
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    double xb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xb[d] = bst[d * Nbp + b] + alpha_end * bst[(3 + d) * Nbp + b];
    const int k = r / kWave;
    const int base = a.slice_base[R.slice_off + k] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const double l0 = gl0[idx];
      const int w = wv & 0xFFFFFF;
      const double dx = P[w] - xb[0];
      const double dy = P[Np + w] - xb[1];
      const double dz = P[2 * Np + w] - xb[2];
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      double frc;
      if (lin_only) {
        double lr = l / l0;
        double gs = 1.0 / 2.0 * (lr * lr - 1);
        frc = lr * lin.E * lin.A * gs;
      } else {
        frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
      }
      const double t = frc / l;
      f1 = f1 - dx * t;
      f2 = f2 - dy * t;
      f3 = f3 - dz * t;
    }
    st[0] += xb[0] * f1;
    st[1] += xb[1] * f2;
    st[2] += xb[2] * f3;
    st[3] += 0.5 * (xb[1] * f3 + xb[2] * f2);
    st[4] += 0.5 * (xb[0] * f3 + xb[2] * f1);
    st[5] += 0.5 * (xb[0] * f2 + xb[1] * f1);
    if (a.solve_flags[sid] & kWriteX) {
      for (int d = 0; d < 3; ++d) a.x_out[3 * (nb + r) + d] = xb[d];
    }
    if (a.solve_flags[sid] & kWriteU) {
      for (int d = 0; d < 3; ++d)
        a.u_out[3 * (nb + r) + d] = xb[d] - a.X[3 * (nb + r) + d];
    }
  }
  // free nodes: u = x - X (TeamOuterLoop.h:427-432)
  const int flags = a.solve_flags[sid];
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      const int64_t g = nb + r;
      if (flags & kWriteX)
        for (int d = 0; d < 3; ++d) a.x_out[3 * g + d] = xr[s][d];
      if (flags & kWriteU)
        for (int d = 0; d < 3; ++d) a.u_out[3 * g + d] = xr[s][d] - a.X[3 * g + d];
    }
  }
  block_sum<B, 6>(st, red);
  if (tid == 0) {
    const double sv = R.scale / a.V[rv];
    bool bad = false;
    for (int i = 0; i < 6; ++i) {
      double v = st[i] * sv;
      a.sigma[6 * (int64_t)sid + i] = v;
      bad = bad || !isfinite(v);
    }
    if (bad && status == MF_STATUS_OK) status = MF_STATUS_NAN;
    a.status[sid] = status;
    a.steps[sid] = step;
  }
}

// ============================================================================
// FP64 kernel, each fiber evaluated once per step
// ============================================================================
// Bitwise the result of solve64: at end b the difference vector is exactly
// the negation of end a's (x_a - x_b = -(x_b - x_a) in IEEE arithmetic), so
// the length, the force and t = f/l of a fiber are the same numbers at both
// ends.  A fiber phase computes t once per fiber with a free end (fiber
// stream order, l0 and reaction from end a's ELL entry) into LDS; each free
// node then recomputes its difference vectors and subtracts d t in its ELL
// order, taking t by the entry's fiber reference, which are solve64's
// operations and order without the second square root and divisions.
// Positions double-buffered in LDS; two barriers per step (positions
// published, t published).
template <int B, int NPT>
__global__ __launch_bounds__(B) void solve64f(SolveArgs a) {
#pragma clang fp contract(off)
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const double eta_r = a.eta_rve ? a.eta_rve[rv] : a.eta;
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M, Mf = R.m_free;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int64_t nb = R.node_off;

  // LDS carve (mirrors lds_bytes)
  size_t off = 0;
  double *pos = reinterpret_cast<double *>(smem);  // [2][3][Np]
  off += sizeof(double) * 6 * Np;
  double *trec = reinterpret_cast<double *>(smem + off);  // [Mf] t per fiber
  off += sizeof(double) * ((Mf + 1) & ~1);
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * 6 * Nbp;
  ReactDev *rxs = reinterpret_cast<ReactDev *>(smem + off);

  const int32_t *gw = a.ell_w + R.ell_off;
  const double *gl0 = a.ell_l0 + R.ell_off;
  const int32_t *gfr = a.ell_fref + R.ell_off;
  const int32_t *fst = a.fstream + 4 * R.fib_off;
  if (tid < R.n_react) rxs[tid] = a.reacts[R.react_off + tid];
  const bool lin_only =
      R.n_react == 1 && a.reacts[R.react_off].type == MF_REACTION_LINEAR;
  const Lin64 lin{a.reacts[R.react_off].E, a.reacts[R.react_off].A};

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;

  // ---- initial state (as solve64) -------------------------------------------
  double xr[NPT][3], vr[NPT][3], fr[NPT][3], mr[NPT];
  int degr[NPT], Dr[NPT], baser[NPT];
  auto init_node = [&](int r, double (&x)[3]) {
    const int64_t g = nb + r;
    double xp[3], up[3], u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = a.x_in[3 * g + d];
      up[d] = a.u_in[3 * g + d];
    }
    double inc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      inc[k] = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
               (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
               (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      u[k] = ramp ? up[k] : inc[k] + up[k];
      x[k] = a.X[3 * g + k] + u[k];
      pos[k * Np + r] = x[k];
      pos[3 * Np + k * Np + r] = x[k];
    }
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        bst[k * Nbp + b] = x[k];
        bst[(3 + k) * Nbp + b] = ramp ? inc[k] : 0.0;
      }
    }
  };
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    double x[3] = {0, 0, 0};
    if (r < N) init_node(r, x);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xr[s][d] = x[d];
      vr[s][d] = 0.0;
      fr[s][d] = 0.0;
    }
    degr[s] = 0;
    Dr[s] = 0;
    baser[s] = 0;
    mr[s] = 1.0;
    if (r < Nf) {
      mr[s] = a.mass[nb + r];
      degr[s] = a.deg[nb + r];
      const int k = r / kWave;
      Dr[s] = a.slice_D[R.slice_off + k];
      baser[s] = a.slice_base[R.slice_off + k] + lane;
    }
  }
  for (int r = NPT * B + tid; r < N; r += B) {
    double x[3];
    init_node(r, x);
  }
  __syncthreads();

  const double c_sound = R.c_sound;
  int cur = 0;
  // t = f/l of every fiber with a free end, at the positions in buffer `buf`,
  // with end a's view of the fiber (d = x_b - x_a)
  auto fiber_phase = [&](int buf) {
    const double *P = pos + buf * 3 * Np;
    for (int e = tid; e < Mf; e += B) {
      const int ra = fst[4 * e], rb = fst[4 * e + 1], sa = fst[4 * e + 2];
      const int32_t wv = gw[sa];
      const double l0 = gl0[sa];
      const double dx = P[rb] - P[ra];
      const double dy = P[Np + rb] - P[Np + ra];
      const double dz = P[2 * Np + rb] - P[2 * Np + ra];
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      double frc;
      if (lin_only) {
        double lr = l / l0;
        double gs = 1.0 / 2.0 * (lr * lr - 1);
        frc = lr * lin.E * lin.A * gs;
      } else {
        frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
      }
      trec[e] = frc / l;
    }
  };
  // internal force of a free node: its fibers in ELL order, t by reference
  auto node_force = [&](int s, int buf, double (&f)[3]) {
    const double *P = pos + buf * 3 * Np;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int D = __builtin_amdgcn_readfirstlane(Dr[s]);
    int32_t wv_n = D > 0 ? gw[baser[s]] : 0;
    int32_t fr_n = D > 0 ? gfr[baser[s]] : 0;
    for (int j = 0; j < D; ++j) {
      const int32_t wv = wv_n, frf = fr_n;
      if (j + 1 < D) {
        wv_n = gw[baser[s] + (j + 1) * kWave];
        fr_n = gfr[baser[s] + (j + 1) * kWave];
      }
      if (j < degr[s]) {
        const int w = wv & 0xFFFFFF;
        const double t = trec[frf & 0x7FFFFFFF];
        const double dx = P[w] - xr[s][0];
        const double dy = P[Np + w] - xr[s][1];
        const double dz = P[2 * Np + w] - xr[s][2];
        fx = fx - dx * t;
        fy = fy - dy * t;
        fz = fz - dz * t;
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  auto min_length = [&](int buf) {
    const double *P = pos + buf * 3 * Np;
    double lmin = 1.0e300;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += B) {
      const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
      double x1 = P[n2] - P[n1];
      double x2 = P[Np + n2] - P[Np + n1];
      double x3 = P[2 * Np + n2] - P[2 * Np + n1];
      double l = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
      lmin = l < lmin ? l : lmin;
    }
    return lmin;
  };

  // ---- RunInitialization (TeamOuterLoop.h:215-289) ---------------------------
  double fint[NPT][3];
  double resid_local = 0.0;
  fiber_phase(cur);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    if (s * B + tid < Nf) {
      node_force(s, cur, fint[s]);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double mass = mr[s] * eta_r;
        double fdamp = mass * vr[s][d];
        fr[s][d] = -fint[s][d] - fdamp;
      }
    }
  }
  double lmin = min_length(cur);
  block_min_sum<B>(lmin, resid_local, red);
  double dt_crit = lmin / c_sound;

  // ---- main loop (TeamOuterLoop.h:340-371) -----------------------------------
  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  for (;;) {
    const double h = a.crit * dt_crit * 0.5;
    const int nxt = cur ^ 1;
    double *Pn = pos + nxt * 3 * Np;
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double vl = vr[s][d] + h * fr[s][d] / mr[s];
          vr[s][d] = vl;
          xr[s][d] = xr[s][d] + 2 * h * vl;
          Pn[d * Np + r] = xr[s][d];
        }
      }
    }
    if (ramp && tid < Nb) {
      const double alpha = ramp_alpha(step + 1, a.n_steps, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          Pn[d * Np + Nf + b] = bst[d * Nbp + b] + alpha * bst[(3 + d) * Nbp + b];
      }
    }
    __syncthreads();  // positions published; every t read of the last step done
    cur = nxt;
    fiber_phase(cur);
    __syncthreads();  // t published
    const bool refresh = (step % freq) == 0;
    double rsum = 0.0;
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
        node_force(s, cur, fint[s]);
        double rn = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double mass = mr[s] * eta_r;
          double fdamp = mass * vr[s][d];
          double lres = -fint[s][d];
          fr[s][d] = lres - fdamp;
          rn += lres * lres;
        }
        rsum += rn;
      }
    }
    if (refresh) {
      double lm = min_length(cur);
      block_min_sum<B>(lm, rsum, red);
      dt_crit = lm / c_sound;
      residual = sqrt(rsum);
    }
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d) vr[s][d] += h * fr[s][d] / mr[s];
      }
    }
    ++step;
    if (ramp) {
      if (step >= a.n_steps) break;
    } else {
      if (!(residual > a.tol)) break;
      if (step >= a.max_steps) {
        status = MF_STATUS_STEP_CAP;
        break;
      }
    }
  }
  if (residual != residual) status = MF_STATUS_NAN;

  // ---- FinalizeData + computeCauchyStress (as solve64) ----------------------
  const double *P = pos + cur * 3 * Np;
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    double xb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xb[d] = bst[d * Nbp + b] + alpha_end * bst[(3 + d) * Nbp + b];
    const int base = a.slice_base[R.slice_off + r / kWave] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const double l0 = gl0[idx];
      const int w = wv & 0xFFFFFF;
      const double dx = P[w] - xb[0];
      const double dy = P[Np + w] - xb[1];
      const double dz = P[2 * Np + w] - xb[2];
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      double frc;
      if (lin_only) {
        double lr = l / l0;
        double gs = 1.0 / 2.0 * (lr * lr - 1);
        frc = lr * lin.E * lin.A * gs;
      } else {
        frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
      }
      const double t = frc / l;
      f1 = f1 - dx * t;
      f2 = f2 - dy * t;
      f3 = f3 - dz * t;
    }
    st[0] += xb[0] * f1;
    st[1] += xb[1] * f2;
    st[2] += xb[2] * f3;
    st[3] += 0.5 * (xb[1] * f3 + xb[2] * f2);
    st[4] += 0.5 * (xb[0] * f3 + xb[2] * f1);
    st[5] += 0.5 * (xb[0] * f2 + xb[1] * f1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * (nb + r) + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * (nb + r) + d] = xb[d] - a.X[3 * (nb + r) + d];
  }
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      const int64_t g = nb + r;
      if (flags & kWriteX)
        for (int d = 0; d < 3; ++d) a.x_out[3 * g + d] = xr[s][d];
      if (flags & kWriteU)
        for (int d = 0; d < 3; ++d) a.u_out[3 * g + d] = xr[s][d] - a.X[3 * g + d];
    }
  }
  block_sum<B, 6>(st, red);
  if (tid == 0) {
    const double sv = R.scale / a.V[rv];
    bool bad = false;
    for (int i = 0; i < 6; ++i) {
      double v = st[i] * sv;
      a.sigma[6 * (int64_t)sid + i] = v;
      bad = bad || !isfinite(v);
    }
    if (bad && status == MF_STATUS_OK) status = MF_STATUS_NAN;
    a.status[sid] = status;
    a.steps[sid] = step;
  }
}

// ============================================================================
// FP64 kernel with the state in global memory (any RVE size)
// ============================================================================
// The RVEs no LDS/register kernel holds (more free nodes than lanes, or
// positions beyond LDS).  Per node the arithmetic and its order are
// solve64's; positions (double-buffered), velocities and forces live in a
// per-solve global scratch, a.gstate + blockIdx.x * a.gstride doubles laid
// out [2][3][Np] positions, [3][Np] velocities, [3][Np] forces, and each
// thread loops over its free nodes r = tid, tid + B, ...  The workgroup
// barrier orders its waves' global accesses (one CU).  LDS keeps the
// reduction scratch, the boundary state and the reaction table.
template <int B>
__global__ __launch_bounds__(B) void solve64g(SolveArgs a) {
#pragma clang fp contract(off)
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const double eta_r = a.eta_rve ? a.eta_rve[rv] : a.eta;
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int64_t nb = R.node_off;

  // LDS carve (mirrors lds_bytes)
  size_t off = 0;
  double *red = reinterpret_cast<double *>(smem);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * 6 * Nbp;
  ReactDev *rxs = reinterpret_cast<ReactDev *>(smem + off);
  // global state of this solve
  double *pos = a.gstate + (size_t)blockIdx.x * (size_t)a.gstride;  // [2][3][Np]
  double *gv = pos + 6 * (size_t)Np;                                // [3][Np]
  double *gf = gv + 3 * (size_t)Np;                                 // [3][Np]

  const int32_t *gw = a.ell_w + R.ell_off;
  const double *gl0 = a.ell_l0 + R.ell_off;
  if (tid < R.n_react) rxs[tid] = a.reacts[R.react_off + tid];
  const bool lin_only =
      R.n_react == 1 && a.reacts[R.react_off].type == MF_REACTION_LINEAR;
  const Lin64 lin{a.reacts[R.react_off].E, a.reacts[R.react_off].A};

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;

  // ---- initial state (as solve64) -------------------------------------------
  for (int r = tid; r < N; r += B) {
    const int64_t g = nb + r;
    double xp[3], up[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = a.x_in[3 * g + d];
      up[d] = a.u_in[3 * g + d];
    }
    double inc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      inc[k] = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
               (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
               (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double u = ramp ? up[k] : inc[k] + up[k];
      const double x = a.X[3 * g + k] + u;
      pos[k * Np + r] = x;
      pos[3 * Np + k * Np + r] = x;
      if (r < Nf) gv[k * Np + r] = 0.0;
    }
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        bst[k * Nbp + b] = pos[k * Np + r];
        bst[(3 + k) * Nbp + b] = ramp ? inc[k] : 0.0;
      }
    }
  }
  __syncthreads();

  const double c_sound = R.c_sound;
  int cur = 0;
  // internal force at node r (free or boundary) from the positions in buffer
  // `buf`: the node's fibers in increasing fiber index
  auto node_force = [&](int r, int buf, double (&f)[3]) {
    const double *P = pos + buf * 3 * Np;
    const double x0 = P[r], x1 = P[Np + r], x2 = P[2 * Np + r];
    const int base = a.slice_base[R.slice_off + r / kWave] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const double l0 = gl0[idx];
      const int w = wv & 0xFFFFFF;
      const double dx = P[w] - x0;
      const double dy = P[Np + w] - x1;
      const double dz = P[2 * Np + w] - x2;
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      double frc;
      if (lin_only) {
        double lr = l / l0;
        double gs = 1.0 / 2.0 * (lr * lr - 1);
        frc = lr * lin.E * lin.A * gs;
      } else {
        frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
      }
      const double t = frc / l;
      fx = fx - dx * t;
      fy = fy - dy * t;
      fz = fz - dz * t;
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  auto min_length = [&](int buf) {
    const double *P = pos + buf * 3 * Np;
    double lmin = 1.0e300;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += B) {
      const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
      double x1 = P[n2] - P[n1];
      double x2 = P[Np + n2] - P[Np + n1];
      double x3 = P[2 * Np + n2] - P[2 * Np + n1];
      double l = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
      lmin = l < lmin ? l : lmin;
    }
    return lmin;
  };

  // ---- RunInitialization (TeamOuterLoop.h:215-289) ---------------------------
  double resid_local = 0.0;
  for (int r = tid; r < Nf; r += B) {
    double fint[3];
    node_force(r, cur, fint);
    const double mr = a.mass[nb + r];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double mass = mr * eta_r;
      double fdamp = mass * gv[d * Np + r];
      gf[d * Np + r] = -fint[d] - fdamp;
    }
  }
  double lmin = min_length(cur);
  block_min_sum<B>(lmin, resid_local, red);
  double dt_crit = lmin / c_sound;

  // ---- main loop (TeamOuterLoop.h:340-371) -----------------------------------
  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  for (;;) {
    const double h = a.crit * dt_crit * 0.5;
    const int nxt = cur ^ 1;
    const double *Pc = pos + cur * 3 * Np;
    double *Pn = pos + nxt * 3 * Np;
    // updates(): half kick + drift of free nodes (TeamOuterLoop.h:157-175)
    for (int r = tid; r < Nf; r += B) {
      const double mr = a.mass[nb + r];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double vl = gv[d * Np + r] + h * gf[d * Np + r] / mr;
        gv[d * Np + r] = vl;
        Pn[d * Np + r] = Pc[d * Np + r] + 2 * h * vl;
      }
    }
    if (ramp && tid < Nb) {
      const double alpha = ramp_alpha(step + 1, a.n_steps, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          Pn[d * Np + Nf + b] = bst[d * Nbp + b] + alpha * bst[(3 + d) * Nbp + b];
      }
    }
    __syncthreads();
    cur = nxt;
    const bool refresh = (step % freq) == 0;
    double rsum = 0.0;
    for (int r = tid; r < Nf; r += B) {
      double fint[3];
      node_force(r, cur, fint);
      const double mr = a.mass[nb + r];
      double rn = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double mass = mr * eta_r;
        double fdamp = mass * gv[d * Np + r];
        double lres = -fint[d];
        gf[d * Np + r] = lres - fdamp;
        rn += lres * lres;
      }
      rsum += rn;
    }
    if (refresh) {
      double lm = min_length(cur);
      block_min_sum<B>(lm, rsum, red);
      dt_crit = lm / c_sound;
      residual = sqrt(rsum);
    }
    // updateAccelVels(): second half kick (TeamOuterLoop.h:144-156)
    for (int r = tid; r < Nf; r += B) {
      const double mr = a.mass[nb + r];
#pragma unroll
      for (int d = 0; d < 3; ++d) gv[d * Np + r] += h * gf[d * Np + r] / mr;
    }
    ++step;
    if (ramp) {
      if (step >= a.n_steps) break;
    } else {
      if (!(residual > a.tol)) break;
      if (step >= a.max_steps) {
        status = MF_STATUS_STEP_CAP;
        break;
      }
    }
  }
  if (residual != residual) status = MF_STATUS_NAN;

  // ---- FinalizeData + computeCauchyStress (as solve64) ----------------------
  const double *P = pos + cur * 3 * Np;
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    double xb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xb[d] = bst[d * Nbp + b] + alpha_end * bst[(3 + d) * Nbp + b];
    const int base = a.slice_base[R.slice_off + r / kWave] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const double l0 = gl0[idx];
      const int w = wv & 0xFFFFFF;
      const double dx = P[w] - xb[0];
      const double dy = P[Np + w] - xb[1];
      const double dz = P[2 * Np + w] - xb[2];
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      double frc;
      if (lin_only) {
        double lr = l / l0;
        double gs = 1.0 / 2.0 * (lr * lr - 1);
        frc = lr * lin.E * lin.A * gs;
      } else {
        frc = react64(rxs[(uint32_t)wv >> 24], l0, l);
      }
      const double t = frc / l;
      f1 = f1 - dx * t;
      f2 = f2 - dy * t;
      f3 = f3 - dz * t;
    }
    st[0] += xb[0] * f1;
    st[1] += xb[1] * f2;
    st[2] += xb[2] * f3;
    st[3] += 0.5 * (xb[1] * f3 + xb[2] * f2);
    st[4] += 0.5 * (xb[0] * f3 + xb[2] * f1);
    st[5] += 0.5 * (xb[0] * f2 + xb[1] * f1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * (nb + r) + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * (nb + r) + d] = xb[d] - a.X[3 * (nb + r) + d];
  }
  // free nodes: u = x - X (TeamOuterLoop.h:427-432)
  for (int r = tid; r < Nf; r += B) {
    const int64_t g = nb + r;
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * g + d] = P[d * Np + r];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * g + d] = P[d * Np + r] - a.X[3 * g + d];
  }
  block_sum<B, 6>(st, red);
  if (tid == 0) {
    const double sv = R.scale / a.V[rv];
    bool bad = false;
    for (int i = 0; i < 6; ++i) {
      double v = st[i] * sv;
      a.sigma[6 * (int64_t)sid + i] = v;
      bad = bad || !isfinite(v);
    }
    if (bad && status == MF_STATUS_OK) status = MF_STATUS_NAN;
    a.status[sid] = status;
    a.steps[sid] = step;
  }
}

}  // namespace kern

hipError_t launch_solve64(const Variant &v, const SolveArgs &a, int n_blocks,
                  size_t lds, hipStream_t s) {
#define MF_VARIANT_LAUNCH(P, B, N, F)                                        \
  if constexpr (P == MF_PREC_FP64)                                           \
    if (v.block == B && v.npt == N && (v.flags & ~kEnergy) == F) {          \
    auto k = (v.flags & kEnergy) ? kern::solve64<B, N, F, true>              \
                                 : kern::solve64<B, N, F, false>;            \
    if (lds > 64 * 1024)                                                     \
      (void)hipFuncSetAttribute((const void *)k,                             \
                                hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)kMaxLds);                               \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);               \
    return hipGetLastError();                                                \
  }
  MF_VARIANT_LIST(MF_VARIANT_LAUNCH)
#undef MF_VARIANT_LAUNCH
#define MF_FONCE_LAUNCH(B, N)                                                 \
  if (v.block == B && v.npt == N && v.flags == kFiberOnce) {                 \
    auto k = kern::solve64f<B, N>;                                           \
    if (lds > 64 * 1024)                                                     \
      (void)hipFuncSetAttribute((const void *)k,                             \
                                hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)kMaxLds);                               \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);               \
    return hipGetLastError();                                                \
  }
  MF_FONCE_LAUNCH(256, 1)
  MF_FONCE_LAUNCH(256, 2)
  MF_FONCE_LAUNCH(512, 1)
  MF_FONCE_LAUNCH(512, 2)
  MF_FONCE_LAUNCH(1024, 1)
  MF_FONCE_LAUNCH(1024, 2)
#undef MF_FONCE_LAUNCH
  if (v.block == 1024 && v.flags == kGlobalState) {
    auto k = kern::solve64g<1024>;
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kMaxLds);
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(1024), lds, s, a);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

}  // namespace mf
