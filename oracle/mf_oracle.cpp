// TEST INFRASTRUCTURE ONLY.  CPU FP64 oracle of the batched explicit
// fiber-network RVE solve.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load this library, and only to check or time
// the HIP path; the product never links or calls it.
//
// It restates the reference algorithm literally, in the reference's node
// order and operation order:
//   constructor        BatchedFiberRVEAnalysisExplicit.h:80-256
//   boundary-last order BatchedReorderMesh.h:54-150 (init, scan, permute)
//   nodal mass         MassIntegrator.cc:16-55 (row-sum lumping of linear
//                      truss shape functions = rho*A*L0/2 per fiber end)
//   run()              BatchedFiberRVEAnalysisExplicit.h:257-320
//   affine increment   BatchedFiberRVEAnalysisExplicitBase.h:16-57
//   volume update      BatchedFiberRVEAnalysisExplicitBase.h:127-154
//   team loop          BatchedFiberRVEAnalysisExplicitTeamOuterLoop.h:180-437
//   element ops        BatchedFiberRVEAnalysisExplicitBase.h:155-287
//   stress             BatchedFiberRVEAnalysisExplicitBase.h:58-126
//   tangent            BatchedFiberRVEAnalysisExplicit.h:321-405
//   orientation        BatchedComputeOrientationTensor.h:56-176
//   reaction laws      FiberReactions.cc:5-71
// plus the ramp mode (fixed steps, boundary ramped linearly from the
// committed state) and the central-difference tangent, which have no
// reference counterpart and are defined here the same way as in the HIP
// path (parity of those parts is pinned by closed forms and invariants, not
// by reference outputs).
//
// Pinning (tests/test_oracle.py): ReorderMesh known-answer test
// (test/microscale/test_reorder_mesh.cc:9-66), Network_2_7 golden stress and
// tangent (test/microscale/before-refactor-outputs/explicit_cpu_out.txt:27-50,
// scaled by 1/det F), single-fiber and chain closed forms, the nodal-mass sum,
// the virial identity, affine patch test, orientation tensor of
// Network_2_7.txt.summary.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/mumfim_amd.h"

namespace {

struct OReaction {
  int type;
  double r, E, rho, B, lt, lc;
  double area() const { return r * r * M_PI; }
};

// Linear truss reaction as used by the batched path
// (getLinearReactionForce, ...Base.h:155-172).
double linear_force(double orig_length, double length, double E, double A) {
  double length_ratio = length / orig_length;
  double green_strain = 1.0 / 2.0 * (length_ratio * length_ratio - 1);
  return length_ratio * E * A * green_strain;
}
double nl_force(double alpha, double B, double lr, double g) {
  return alpha * lr * (exp(B * g) - 1);
}
double nl_deriv(double alpha, double B, double lr, double l0, double g) {
  return alpha / l0 * ((1 + lr * lr * B) * exp(B * g) - 1);
}
// NonlinearReaction::forceReaction (FiberReactions.cc:32-71), force only.
double nonlinear_force(const OReaction &q, double orig_length, double length) {
  double length_ratio = length / orig_length;
  double alpha = q.E * q.area() / q.B;
  if (length_ratio > q.lt) {
    double gt = 1.0 / 2.0 * (q.lt * q.lt - 1);
    double m = nl_deriv(alpha, q.B, q.lt, orig_length, gt);
    double b = nl_force(alpha, q.B, q.lt, gt);
    return m * (length_ratio - q.lt) * orig_length + b;
  } else if (length_ratio < q.lc) {
    double gc = 1.0 / 2.0 * (q.lc * q.lc - 1);
    double m = nl_deriv(alpha, q.B, q.lc, orig_length, gc);
    double b = nl_force(alpha, q.B, q.lc, gc);
    return m * (length_ratio - q.lc) * orig_length + b;
  }
  double g = 1.0 / 2.0 * (length_ratio * length_ratio - 1);
  return nl_force(alpha, q.B, length_ratio, g);
}

struct ORve {
  int N = 0, M = 0, Nb = 0;
  std::vector<int> conn;       // 2M, reference order
  std::vector<double> X;       // 3N original coordinates (reordered)
  std::vector<double> x, u, v, f, f_int, mass, l0, l;
  std::vector<int> fixed;      // boundary verts after reorder
  std::vector<int> perm;       // input node -> reference index
  std::vector<int> fib_react;  // M
  std::vector<OReaction> reactions;
  double E = 0, A = 0, rho = 0;  // reaction 0 (batched path)
  double c = 0;                  // sound speed used for dt
  double V = 1.0, scale = 1.0;
  bool per_fiber = false;
  // ramp-mode state before the last update run
  std::vector<double> x_prev, u_prev;
  double F_last[9];
  double current_stress[6] = {0, 0, 0, 0, 0, 0};
  // W_int, W_ext, W_damp, W_kin of the last solve (energy diagnostics)
  double energy[4] = {0, 0, 0, 0};
};

struct OBatch {
  mf_solver_params p;
  std::vector<ORve> rves;
  bool have_prev = false;
};

// ReorderMesh::createPermutationArray (BatchedReorderMesh.h:53-73,101-110):
// perm(fixed[i]) = i + 1, then an inclusive scan of "is fixed" gives
// free i -> i - count, fixed i -> (N - Nb) + (perm - 1).
std::vector<int> reorder_permutation(int nvert, const std::vector<int> &fixed) {
  std::vector<int> perm(nvert, 0);
  for (size_t i = 0; i < fixed.size(); ++i) perm[fixed[i]] = (int)i + 1;
  double update = 0;  // the reference scans with a Scalar accumulator
  const int nfixed = (int)fixed.size();
  for (int i = 0; i < nvert; ++i) {
    int updateVal = perm[i];
    bool isFixedVert = updateVal > 0;
    update += isFixedVert ? 1 : 0;
    perm[i] = isFixedVert ? (nvert - nfixed) + (updateVal - 1) : i - (int)update;
  }
  return perm;
}

double fiber_force(const ORve &R, int e, double l0, double l) {
  if (!R.per_fiber) return linear_force(l0, l, R.E, R.A);
  const OReaction &q = R.reactions[R.fib_react[e]];
  if (q.type == MF_REACTION_LINEAR) return linear_force(l0, l, q.E, q.area());
  return nonlinear_force(q, l0, l);
}

// getElementLength (...Base.h:203-215) over all fibers; returns min length.
double element_lengths(const ORve &R, const std::vector<double> &c,
                       std::vector<double> &out) {
  double lmin = 1e300;
  for (int i = 0; i < R.M; ++i) {
    int n1 = R.conn[2 * i], n2 = R.conn[2 * i + 1];
    double x1 = c[3 * n2] - c[3 * n1];
    double x2 = c[3 * n2 + 1] - c[3 * n1 + 1];
    double x3 = c[3 * n2 + 2] - c[3 * n1 + 2];
    out[i] = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    lmin = out[i] < lmin ? out[i] : lmin;
  }
  return lmin;
}

// getForces (TeamOuterLoop.h:83-143) over vertices [0, nv); returns the
// residual sqrt(sum f_int^2) of those vertices.
double get_forces(ORve &R, int nv, double eta) {
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < 3; ++j) R.f_int[3 * i + j] = 0;
  for (int i = 0; i < R.M; ++i) {
    double local_l = R.l[i];
    double frc = fiber_force(R, i, R.l0[i], local_l);
    double frc_ovr_l = frc / local_l;
    int n1 = R.conn[2 * i], n2 = R.conn[2 * i + 1];
    double e1 = (R.x[3 * n2] - R.x[3 * n1]) * frc_ovr_l;
    double e2 = (R.x[3 * n2 + 1] - R.x[3 * n1 + 1]) * frc_ovr_l;
    double e3 = (R.x[3 * n2 + 2] - R.x[3 * n1 + 2]) * frc_ovr_l;
    R.f_int[3 * n1] -= e1;
    R.f_int[3 * n2] += e1;
    R.f_int[3 * n1 + 1] -= e2;
    R.f_int[3 * n2 + 1] += e2;
    R.f_int[3 * n1 + 2] -= e3;
    R.f_int[3 * n2 + 2] += e3;
  }
  double residual = 0;
  for (int i = 0; i < nv; ++i) {
    double mass = R.mass[i] * eta;
    double r = 0;
    for (int j = 0; j < 3; ++j) {
      double f_damp = mass * R.v[3 * i + j];
      double local_residual = -R.f_int[3 * i + j];
      R.f[3 * i + j] = local_residual - f_damp;
      r += local_residual * local_residual;
    }
    residual += r;
  }
  return sqrt(residual);
}

// Ramp amplitude after step k of S: linear, or SmoothAmp
// (ExplicitAmplitude.h:26-40).
double ramp_alpha(int k, int S, int amplitude) {
  const double xi = (double)k / (double)S;
  if (amplitude == MF_AMP_SMOOTHSTEP) return xi * xi * xi * (10 - 15 * xi + 6 * xi * xi);
  return xi;
}

// Internal forces of the boundary vertices (indices Nf..N-1) by a scatter
// over every fiber in fiber order (the getForces scatter, restricted).
void boundary_forces(const ORve &R, std::vector<double> &fb) {
  const int Nf = R.N - R.Nb;
  std::fill(fb.begin(), fb.end(), 0.0);
  for (int i = 0; i < R.M; ++i) {
    const int n1 = R.conn[2 * i], n2 = R.conn[2 * i + 1];
    if (n1 < Nf && n2 < Nf) continue;
    const double frc = fiber_force(R, i, R.l0[i], R.l[i]);
    const double frc_ovr_l = frc / R.l[i];
    for (int k = 0; k < 3; ++k) {
      const double e = (R.x[3 * n2 + k] - R.x[3 * n1 + k]) * frc_ovr_l;
      if (n1 >= Nf) fb[3 * (n1 - Nf) + k] -= e;
      if (n2 >= Nf) fb[3 * (n2 - Nf) + k] += e;
    }
  }
}

// checkEnergyBalance_ (FiberRVEAnalysisExplicit_impl.h:590-600)
bool energy_balance_ok(const double W[4], double eps) {
  const double bal = W[3] + W[0] + W[2] - W[1];
  return fabs(bal) <= eps * std::max(std::max(fabs(W[1]), std::max(fabs(W[0]), fabs(W[3]))),
                                     fabs(W[2]));
}

// One solve of one RVE.  Returns steps taken; status in *status.
int solve_rve(ORve &R, const mf_solver_params &p, const double F[9],
              bool update, double sigma[6], int *status) {
  const int N = R.N, Nf = R.N - R.Nb;
  const bool ramp = p.mode == MF_MODE_RAMP;
  std::vector<double> u_copy, x_copy;
  if (!update) {
    u_copy = R.u;
    if (ramp) x_copy = R.x;
  } else {
    // updateVolume (...Base.h:143-153)
    const double *dg = F;
    double J = dg[0] * (dg[4] * dg[8] - dg[5] * dg[7]) -
               dg[1] * (dg[3] * dg[8] - dg[5] * dg[6]) +
               dg[2] * (dg[3] * dg[7] - dg[4] * dg[6]);
    R.V = J * R.V;
    if (ramp) {
      R.x_prev = R.x;
      R.u_prev = R.u;
      std::memcpy(R.F_last, F, sizeof(R.F_last));
    }
  }
  // affine increment (F - I) x of the current coordinates
  std::vector<double> g(3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < 3; ++k)
      g[3 * i + k] = (F[3 * k + 0] - (k == 0 ? 1 : 0)) * R.x[3 * i] +
                     (F[3 * k + 1] - (k == 1 ? 1 : 0)) * R.x[3 * i + 1] +
                     (F[3 * k + 2] - (k == 2 ? 1 : 0)) * R.x[3 * i + 2];
  if (!ramp)
    for (int i = 0; i < 3 * N; ++i) R.u[i] = g[i] + R.u[i];
  std::fill(R.v.begin(), R.v.end(), 0.0);
  // RunInitialization (TeamOuterLoop.h:215-289)
  double dt_crit = element_lengths(R, R.X, R.l0) / R.c;
  for (int i = 0; i < 3 * N; ++i) R.x[i] = R.X[i] + R.u[i];
  std::vector<double> xb0(3 * (size_t)R.Nb), gb(3 * (size_t)R.Nb);
  for (int b = 0; b < R.Nb; ++b)
    for (int k = 0; k < 3; ++k) {
      xb0[3 * b + k] = R.x[3 * (Nf + b) + k];
      gb[3 * b + k] = ramp ? g[3 * (Nf + b) + k] : 0.0;
    }
  dt_crit = element_lengths(R, R.x, R.l) / R.c;
  get_forces(R, Nf, p.visc_damp_coeff);
  // energy diagnostics (computeEnergies_, computeKineticEnergy_,
  // FiberRVEAnalysisExplicit_impl.h:1101-1138): sums of du.(f_prev + f)
  const bool en = p.energy_diagnostics != 0;
  double w_int = 0, w_ext = 0, w_damp = 0;
  std::vector<double> fl, dl, du, fbl, fbn;
  if (en) {
    fl.assign(R.f_int.begin(), R.f_int.begin() + 3 * (size_t)Nf);
    dl.assign(3 * (size_t)Nf, 0.0);
    du.assign(3 * (size_t)Nf, 0.0);
    fbl.assign(3 * (size_t)R.Nb, 0.0);
    fbn.assign(3 * (size_t)R.Nb, 0.0);
    boundary_forces(R, fbl);
  }
  auto totals = [&](double W[4]) {
    double wk = 0;
    for (int i = 0; i < Nf; ++i)
      for (int j = 0; j < 3; ++j) wk += R.mass[i] * R.v[3 * i + j] * R.v[3 * i + j];
    W[0] = 0.5 * w_int;
    W[1] = 0.5 * w_ext;
    W[2] = 0.5 * w_damp;
    W[3] = 0.5 * wk;
  };
  // main loop (TeamOuterLoop.h:340-371)
  int step = 0;
  double residual = 10.0;
  int st = MF_STATUS_OK;
  const double crit = p.crit_time_scale_factor;
  do {
    double dt = crit * dt_crit * 0.5;
    for (int i = 0; i < Nf; ++i) {
      double mass = R.mass[i];
      for (int j = 0; j < 3; ++j) {
        double v_local = R.v[3 * i + j] + dt * R.f[3 * i + j] / mass;
        R.v[3 * i + j] = v_local;
        R.x[3 * i + j] = R.x[3 * i + j] + 2 * dt * v_local;
        if (en) du[3 * i + j] = 2 * dt * v_local;
      }
    }
    if (ramp) {
      double alpha = ramp_alpha(step + 1, p.n_steps, p.amplitude);
      for (int b = 0; b < R.Nb; ++b)
        for (int k = 0; k < 3; ++k)
          R.x[3 * (Nf + b) + k] = xb0[3 * b + k] + alpha * gb[3 * b + k];
    }
    const bool refresh = step % p.dt_update_freq == 0;
    double lmin = element_lengths(R, R.x, R.l);
    if (refresh) dt_crit = lmin / R.c;
    double res = get_forces(R, Nf, p.visc_damp_coeff);
    if (refresh) residual = res;
    if (en) {
      for (int i = 0; i < Nf; ++i) {
        const double mass = R.mass[i] * p.visc_damp_coeff;
        for (int j = 0; j < 3; ++j) {
          const int q = 3 * i + j;
          const double fdamp = mass * R.v[q];
          w_int += du[q] * (fl[q] + R.f_int[q]);
          w_damp += du[q] * (dl[q] + fdamp);
          fl[q] = R.f_int[q];
          dl[q] = fdamp;
        }
      }
      if (ramp) {
        const double a0 = step == 0 ? 0.0 : ramp_alpha(step, p.n_steps, p.amplitude);
        const double a1 = ramp_alpha(step + 1, p.n_steps, p.amplitude);
        boundary_forces(R, fbn);
        for (int b = 0; b < R.Nb; ++b)
          for (int k = 0; k < 3; ++k) {
            const double x1 = xb0[3 * b + k] + a1 * gb[3 * b + k];
            const double x0 = xb0[3 * b + k] + a0 * gb[3 * b + k];
            const double t = (x1 - x0) * (fbl[3 * b + k] + fbn[3 * b + k]);
            w_int += t;
            w_ext += t;
            fbl[3 * b + k] = fbn[3 * b + k];
          }
      }
    }
    for (int i = 0; i < Nf; ++i) {
      double mass = R.mass[i];
      for (int j = 0; j < 3; ++j) R.v[3 * i + j] += dt * R.f[3 * i + j] / mass;
    }
    ++step;
    if (en && p.energy_check_eps > 0 && step > 10 && step % p.dt_update_freq == 0) {
      double W[4];
      totals(W);
      if (!energy_balance_ok(W, p.energy_check_eps)) {
        st = MF_STATUS_ENERGY;
        break;
      }
    }
    if (ramp) {
      if (step >= p.n_steps) break;
    } else {
      if (!(residual > p.tolerance)) break;
      if (step >= p.max_steps) {
        st = MF_STATUS_STEP_CAP;
        break;
      }
    }
  } while (true);
  if (residual != residual) st = MF_STATUS_NAN;
  if (en) {
    totals(R.energy);
    if (st == MF_STATUS_OK && p.energy_check_eps > 0 && step > 10 &&
        !energy_balance_ok(R.energy, p.energy_check_eps))
      st = MF_STATUS_ENERGY;
  }
  // FinalizeData (TeamOuterLoop.h:374-434)
  get_forces(R, N, p.visc_damp_coeff);
  for (int i = 0; i < 3 * N; ++i) R.u[i] = R.x[i] - R.X[i];
  // computeCauchyStress (...Base.h:84-123)
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < R.Nb; ++b) {
    int vert = R.fixed[b];
    double f1 = R.f_int[3 * vert], f2 = R.f_int[3 * vert + 1],
           f3 = R.f_int[3 * vert + 2];
    double c1 = R.x[3 * vert], c2 = R.x[3 * vert + 1], c3 = R.x[3 * vert + 2];
    s[0] += c1 * f1;
    s[1] += c2 * f2;
    s[2] += c3 * f3;
    s[3] += 0.5 * (c2 * f3 + c3 * f2);
    s[4] += 0.5 * (c1 * f3 + c3 * f1);
    s[5] += 0.5 * (c1 * f2 + c2 * f1);
  }
  for (int i = 0; i < 6; ++i) {
    s[i] *= R.scale / R.V;
    sigma[i] = s[i];
    if (!std::isfinite(s[i]) && st == MF_STATUS_OK) st = MF_STATUS_NAN;
  }
  if (!update) {
    R.u = u_copy;
    if (ramp) R.x = x_copy;
  } else {
    for (int i = 0; i < 6; ++i) R.current_stress[i] = sigma[i];
  }
  if (status) *status = st;
  return step;
}

// ---- tangent-linear (ramp mode) -----------------------------------------
// d f / d l of a fiber's reaction (FiberReactions.cc:5-71): linear
// EA (3 lambda^2 - 1) / (2 l0); toe law alpha/l0 ((1 + lambda^2 B) e^{B g} - 1);
// beyond a transition the slope at the transition.
double fiber_dforce(const ORve &R, int e, double l0, double l) {
  double E = R.E, A = R.A;
  if (R.per_fiber) {
    const OReaction &q = R.reactions[R.fib_react[e]];
    E = q.E;
    A = q.area();
    if (q.type != MF_REACTION_LINEAR) {
      const double lr = l / l0, alpha = E * A / q.B;
      double lam = lr;
      if (lr > q.lt) lam = q.lt;
      else if (lr < q.lc) lam = q.lc;
      return nl_deriv(alpha, q.B, lam, l0, 0.5 * (lam * lam - 1));
    }
  }
  const double lr = l / l0;
  return E * A * (3 * lr * lr - 1) / (2 * l0);
}

// Fiber lengths and their derivatives along dx; returns the minimum length
// (first index on ties, as element_lengths) and its derivative in *dmin.
double element_lengths_tl(const ORve &R, const std::vector<double> &c,
                          const std::vector<double> &dc, std::vector<double> &out,
                          std::vector<double> &dout, double *dmin) {
  double lmin = 1e300, dl = 0;
  for (int i = 0; i < R.M; ++i) {
    const int n1 = R.conn[2 * i], n2 = R.conn[2 * i + 1];
    double s = 0, ds = 0;
    for (int k = 0; k < 3; ++k) {
      const double d = c[3 * n2 + k] - c[3 * n1 + k];
      s += d * d;
      ds += d * (dc[3 * n2 + k] - dc[3 * n1 + k]);
    }
    out[i] = sqrt(s);
    dout[i] = ds / out[i];
    if (out[i] < lmin) {
      lmin = out[i];
      dl = dout[i];
    }
  }
  *dmin = dl;
  return lmin;
}

// Internal forces and their derivatives over all fibers, then the total
// force (and derivative) of the free vertices [0, nv).
void get_forces_tl(ORve &R, int nv, double eta, const std::vector<double> &dx,
                   const std::vector<double> &dl, const std::vector<double> &dv,
                   std::vector<double> &df_int, std::vector<double> &df) {
  std::fill(R.f_int.begin(), R.f_int.end(), 0.0);
  std::fill(df_int.begin(), df_int.end(), 0.0);
  for (int i = 0; i < R.M; ++i) {
    const double l = R.l[i], l0 = R.l0[i];
    const double frc = fiber_force(R, i, l0, l);
    const double t = frc / l;
    const double dt = fiber_dforce(R, i, l0, l) * dl[i] / l - t * dl[i] / l;
    const int n1 = R.conn[2 * i], n2 = R.conn[2 * i + 1];
    for (int k = 0; k < 3; ++k) {
      const double d = R.x[3 * n2 + k] - R.x[3 * n1 + k];
      const double dd = dx[3 * n2 + k] - dx[3 * n1 + k];
      const double e = d * t, de = dd * t + d * dt;
      R.f_int[3 * n1 + k] -= e;
      R.f_int[3 * n2 + k] += e;
      df_int[3 * n1 + k] -= de;
      df_int[3 * n2 + k] += de;
    }
  }
  for (int i = 0; i < nv; ++i) {
    const double mass = R.mass[i] * eta;
    for (int j = 0; j < 3; ++j) {
      R.f[3 * i + j] = -R.f_int[3 * i + j] - mass * R.v[3 * i + j];
      df[3 * i + j] = -df_int[3 * i + j] - mass * dv[3 * i + j];
    }
  }
}

// Tangent-linear of one ramp-mode trial solve from the current state: the
// derivative of the solve's stress along a perturbation of its deformation
// gradient, dF = G (here E_i F_last).  Every quantity the solve computes is
// differentiated in the same order: the boundary ramp (d x_b = alpha G x_b),
// the half kicks and drifts, the fiber forces (d(f/l) from d f/d l), the
// critical step (through the shortest fiber, refreshed with it) and the
// boundary stress sum.  The state is left untouched.
void solve_rve_tl(ORve &R, const mf_solver_params &p, const double F[9],
                  const double G[9], double sigma[6], double dsigma[6]) {
  const int N = R.N, Nf = R.N - R.Nb;
  const std::vector<double> u_copy = R.u, x_copy = R.x;
  std::vector<double> g(3 * (size_t)N), dg(3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < 3; ++k) {
      g[3 * i + k] = (F[3 * k + 0] - (k == 0 ? 1 : 0)) * R.x[3 * i] +
                     (F[3 * k + 1] - (k == 1 ? 1 : 0)) * R.x[3 * i + 1] +
                     (F[3 * k + 2] - (k == 2 ? 1 : 0)) * R.x[3 * i + 2];
      dg[3 * i + k] = G[3 * k + 0] * R.x[3 * i] + G[3 * k + 1] * R.x[3 * i + 1] +
                      G[3 * k + 2] * R.x[3 * i + 2];
    }
  std::fill(R.v.begin(), R.v.end(), 0.0);
  std::vector<double> dx(3 * (size_t)N, 0.0), dv(3 * (size_t)N, 0.0), df(3 * (size_t)N, 0.0),
      df_int(3 * (size_t)N, 0.0), dl(R.M, 0.0);
  double ddt_crit = 0;
  for (int i = 0; i < 3 * N; ++i) R.x[i] = R.X[i] + R.u[i];
  std::vector<double> xb0(3 * (size_t)R.Nb);
  for (int b = 0; b < R.Nb; ++b)
    for (int k = 0; k < 3; ++k) xb0[3 * b + k] = R.x[3 * (Nf + b) + k];
  double dt_crit = element_lengths_tl(R, R.x, dx, R.l, dl, &ddt_crit) / R.c;
  ddt_crit /= R.c;
  get_forces_tl(R, Nf, p.visc_damp_coeff, dx, dl, dv, df_int, df);
  const double crit = p.crit_time_scale_factor;
  int step = 0;
  do {
    const double dt = crit * dt_crit * 0.5, ddt = crit * ddt_crit * 0.5;
    for (int i = 0; i < Nf; ++i) {
      const double mass = R.mass[i];
      for (int j = 0; j < 3; ++j) {
        const int q = 3 * i + j;
        const double v_local = R.v[q] + dt * R.f[q] / mass;
        const double dv_local = dv[q] + ddt * R.f[q] / mass + dt * df[q] / mass;
        R.v[q] = v_local;
        dv[q] = dv_local;
        R.x[q] = R.x[q] + 2 * dt * v_local;
        dx[q] = dx[q] + 2 * (ddt * v_local + dt * dv_local);
      }
    }
    const double alpha = ramp_alpha(step + 1, p.n_steps, p.amplitude);
    for (int b = 0; b < R.Nb; ++b)
      for (int k = 0; k < 3; ++k) {
        const int q = 3 * (Nf + b) + k;
        R.x[q] = xb0[3 * b + k] + alpha * g[q];
        dx[q] = alpha * dg[q];
      }
    double dlmin = 0;
    const double lmin = element_lengths_tl(R, R.x, dx, R.l, dl, &dlmin);
    if (step % p.dt_update_freq == 0) {
      dt_crit = lmin / R.c;
      ddt_crit = dlmin / R.c;
    }
    get_forces_tl(R, Nf, p.visc_damp_coeff, dx, dl, dv, df_int, df);
    for (int i = 0; i < Nf; ++i) {
      const double mass = R.mass[i];
      for (int j = 0; j < 3; ++j) {
        const int q = 3 * i + j;
        R.v[q] += dt * R.f[q] / mass;
        dv[q] += ddt * R.f[q] / mass + dt * df[q] / mass;
      }
    }
    ++step;
  } while (step < p.n_steps);
  get_forces_tl(R, N, p.visc_damp_coeff, dx, dl, dv, df_int, df);
  double s[6] = {0, 0, 0, 0, 0, 0}, ds[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < R.Nb; ++b) {
    const int vert = R.fixed[b];
    const double *f = &R.f_int[3 * vert], *c = &R.x[3 * vert];
    const double *dfi = &df_int[3 * vert], *dc = &dx[3 * vert];
    s[0] += c[0] * f[0];
    s[1] += c[1] * f[1];
    s[2] += c[2] * f[2];
    s[3] += 0.5 * (c[1] * f[2] + c[2] * f[1]);
    s[4] += 0.5 * (c[0] * f[2] + c[2] * f[0]);
    s[5] += 0.5 * (c[0] * f[1] + c[1] * f[0]);
    ds[0] += dc[0] * f[0] + c[0] * dfi[0];
    ds[1] += dc[1] * f[1] + c[1] * dfi[1];
    ds[2] += dc[2] * f[2] + c[2] * dfi[2];
    ds[3] += 0.5 * (dc[1] * f[2] + c[1] * dfi[2] + dc[2] * f[1] + c[2] * dfi[1]);
    ds[4] += 0.5 * (dc[0] * f[2] + c[0] * dfi[2] + dc[2] * f[0] + c[2] * dfi[0]);
    ds[5] += 0.5 * (dc[0] * f[1] + c[0] * dfi[1] + dc[1] * f[0] + c[1] * dfi[0]);
  }
  for (int i = 0; i < 6; ++i) {
    sigma[i] = s[i] * (R.scale / R.V);
    dsigma[i] = ds[i] * (R.scale / R.V);
  }
  R.u = u_copy;
  R.x = x_copy;
}

// d P_i / d h at h = 0 of the perturbation below: the unit Voigt strain of
// direction i with engineering shear (1 on the diagonal, 1/2 off it).
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

// Almansi perturbation (BatchedFiberRVEAnalysisExplicit.h:340-389).
void perturbation(int i, double h, double Fp[9]) {
  double D1 = sqrt(1.0 / (1 - 2 * h));
  const double l1 = 1;
  const double l2 = (1 + h) / (1 - h * h);
  const double l3 = (1 - h) / (1 - h * h);
  double l2pl3 = 0.5 * (sqrt(l2) + sqrt(l3));
  double l2ml3 = 0.5 * (sqrt(l2) - sqrt(l3));
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

}  // namespace

extern "C" {

typedef struct OBatch oracle_batch;

// Packs the networks exactly as the reference constructor does.
int oracle_create(const mf_network *nets, int n, const mf_solver_params *p,
                  oracle_batch **out) {
  OBatch *b = new OBatch();
  b->p = *p;
  b->rves.resize(n);
  for (int r = 0; r < n; ++r) {
    const mf_network &d = nets[r];
    ORve &R = b->rves[r];
    R.N = d.n_nodes;
    R.M = d.n_fibers;
    R.Nb = d.n_boundary;
    R.scale = d.scale_factor;
    R.per_fiber = p->reaction_policy == MF_REACT_PER_FIBER;
    for (int i = 0; i < d.n_reactions; ++i) {
      const mf_reaction &q = d.reactions[i];
      R.reactions.push_back({q.type, q.radius, q.E, q.density, q.B,
                             q.length_ratio_trns, q.cmp_ratio_trns});
    }
    R.E = R.reactions[0].E;
    R.A = R.reactions[0].area();
    R.rho = R.reactions[0].rho;
    R.fib_react.assign(R.M, 0);
    if (R.per_fiber && d.fiber_reaction)
      for (int e = 0; e < R.M; ++e) R.fib_react[e] = d.fiber_reaction[e];
    if (R.per_fiber) {
      R.c = 0;
      std::vector<char> used(R.reactions.size(), 0);
      for (int e = 0; e < R.M; ++e) used[R.fib_react[e]] = 1;
      for (size_t i = 0; i < R.reactions.size(); ++i)
        if (used[i] || R.M == 0) {
          double c = sqrt(R.reactions[i].E / R.reactions[i].rho);
          R.c = c > R.c ? c : R.c;
        }
    } else {
      R.c = sqrt(R.E / R.rho);
    }
    std::vector<int> fixed(d.boundary, d.boundary + R.Nb);
    R.perm = reorder_permutation(R.N, fixed);
    R.X.assign(3 * (size_t)R.N, 0);
    for (int i = 0; i < R.N; ++i)
      for (int j = 0; j < 3; ++j) R.X[3 * R.perm[i] + j] = d.coords[3 * i + j];
    R.conn.resize(2 * (size_t)R.M);
    for (int e = 0; e < 2 * R.M; ++e) R.conn[e] = R.perm[d.fibers[e]];
    R.fixed.resize(R.Nb);
    for (int i = 0; i < R.Nb; ++i) R.fixed[i] = R.perm[fixed[i]];
    R.l0.assign(R.M, 0);
    R.l.assign(R.M, 0);
    element_lengths(R, R.X, R.l0);
    // nodal mass: rho*A*L0/2 to both ends, fibers in order
    R.mass.assign(R.N, 0.0);
    for (int e = 0; e < R.M; ++e) {
      double rho = R.rho, A = R.A;
      if (R.per_fiber) {
        rho = R.reactions[R.fib_react[e]].rho;
        A = R.reactions[R.fib_react[e]].area();
      }
      double c = rho * A * R.l0[e] * 0.5;
      R.mass[R.conn[2 * e]] += c;
      R.mass[R.conn[2 * e + 1]] += c;
    }
    R.x = R.X;
    R.u.assign(3 * (size_t)R.N, 0);
    R.v.assign(3 * (size_t)R.N, 0);
    R.f.assign(3 * (size_t)R.N, 0);
    R.f_int.assign(3 * (size_t)R.N, 0);
  }
  *out = b;
  return 0;
}

void oracle_destroy(oracle_batch *b) { delete b; }

int oracle_set_params(oracle_batch *b, const mf_solver_params *p) {
  b->p = *p;
  return 0;
}

// run(F, sigma, update): one solve per RVE, OpenMP over RVEs.
double oracle_ramp_alpha(int k, int S, int amplitude) { return ramp_alpha(k, S, amplitude); }

int oracle_last_energies(oracle_batch *b, double *W) {
  for (size_t r = 0; r < b->rves.size(); ++r)
    for (int i = 0; i < 4; ++i) W[4 * r + i] = b->rves[r].energy[i];
  return 0;
}

int oracle_run(oracle_batch *b, const double *F, int nF, double *sigma,
               int update, int32_t *status, int32_t *steps) {
  const int n = (int)b->rves.size();
#pragma omp parallel for schedule(dynamic)
  for (int r = 0; r < n; ++r) {
    int st = 0;
    int k = solve_rve(b->rves[r], b->p, F + 9 * (nF == 1 ? 0 : r), update != 0,
                      sigma + 6 * r, &st);
    if (status) status[r] = st;
    if (steps) steps[r] = k;
  }
  return 0;
}

// computeMaterialStiffness: relax mode runs the reference sequence of trial
// runs; ramp mode solves every perturbation from the state before the last
// update run.
int oracle_tangent(oracle_batch *b, double *C) {
  const int n = (int)b->rves.size();
  const double h = b->p.fd_step;
  const bool central = b->p.tangent_scheme == MF_TANGENT_CENTRAL;
#pragma omp parallel for schedule(dynamic)
  for (int r = 0; r < n; ++r) {
    ORve &R = b->rves[r];
    double sp[6], sm[6];
    if (b->p.tangent_scheme == MF_TANGENT_LINEAR) {
      // ramp mode: derivative of each trial solve from the state before the
      // last update run along G = E_i F_last
      std::vector<double> xs = R.x, us = R.u;
      for (int i = 0; i < 6; ++i) {
        double E[9], G[9];
        perturbation_rate(i, E);
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) {
            double s1 = 0;
            for (int k = 0; k < 3; ++k) s1 += E[3 * a + k] * R.F_last[3 * k + c];
            G[3 * a + c] = s1;
          }
        R.x = R.x_prev;
        R.u = R.u_prev;
        solve_rve_tl(R, b->p, R.F_last, G, sp, sm);
        for (int j = 0; j < 6; ++j) C[36 * (size_t)r + 6 * j + i] = sm[j];
      }
      R.x = xs;
      R.u = us;
      continue;
    }
    for (int i = 0; i < 6; ++i) {
      double Fp[9], Fm[9];
      perturbation(i, h, Fp);
      perturbation(i, -h, Fm);
      if (b->p.mode == MF_MODE_RELAX) {
        solve_rve(R, b->p, Fp, false, sp, nullptr);
        if (central) solve_rve(R, b->p, Fm, false, sm, nullptr);
      } else {
        double Ft[9], Ftm[9];
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) {
            double s1 = 0, s2 = 0;
            for (int k = 0; k < 3; ++k) {
              s1 += Fp[3 * a + k] * R.F_last[3 * k + c];
              s2 += Fm[3 * a + k] * R.F_last[3 * k + c];
            }
            Ft[3 * a + c] = s1;
            Ftm[3 * a + c] = s2;
          }
        std::vector<double> xs = R.x, us = R.u;
        R.x = R.x_prev;
        R.u = R.u_prev;
        solve_rve(R, b->p, Ft, false, sp, nullptr);
        if (central) solve_rve(R, b->p, Ftm, false, sm, nullptr);
        R.x = xs;
        R.u = us;
      }
      for (int j = 0; j < 6; ++j)
        C[36 * (size_t)r + 6 * j + i] =
            central ? (sp[j] - sm[j]) / (2 * h) : (sp[j] - R.current_stress[j]) / h;
    }
  }
  return 0;
}

// State in input node order.
int oracle_get_state(oracle_batch *b, double *x, double *u, double *V) {
  size_t off = 0;
  for (size_t r = 0; r < b->rves.size(); ++r) {
    ORve &R = b->rves[r];
    for (int i = 0; i < R.N; ++i)
      for (int j = 0; j < 3; ++j) {
        if (x) x[3 * (off + i) + j] = R.x[3 * R.perm[i] + j];
        if (u) u[3 * (off + i) + j] = R.u[3 * R.perm[i] + j];
      }
    if (V) V[r] = R.V;
    off += R.N;
  }
  return 0;
}

int oracle_orientation(oracle_batch *b, const double *normal, int n_normal,
                       double *omega) {
  const int n = (int)b->rves.size();
  for (int r = 0; r < n; ++r) {
    ORve &R = b->rves[r];
    double nrm[3] = {0, 0, 0};
    if (normal) {
      const double *q = normal + 3 * (n_normal == 1 ? 0 : r);
      double nn = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
      for (int k = 0; k < 3; ++k) nrm[k] = q[k] / nn;
    }
    double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = 0; e < R.M; ++e) {
      int n1 = R.conn[2 * e], n2 = R.conn[2 * e + 1];
      double lx = R.x[3 * n2] - R.x[3 * n1];
      double ly = R.x[3 * n2 + 1] - R.x[3 * n1 + 1];
      double lz = R.x[3 * n2 + 2] - R.x[3 * n1 + 2];
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
      double v[3] = {lx, ly, lz};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[3 * i + j] += v[i] * v[j];
    }
    for (int i = 0; i < 9; ++i) omega[9 * r + i] = S[i] / R.M;
  }
  return 0;
}

// Helpers exposed for known-answer tests.
int oracle_reorder_permutation(int nvert, const int32_t *fixed, int nfixed,
                               int32_t *perm_out) {
  std::vector<int> f(fixed, fixed + nfixed);
  std::vector<int> p = reorder_permutation(nvert, f);
  for (int i = 0; i < nvert; ++i) perm_out[i] = p[i];
  return 0;
}
double oracle_fiber_force(const mf_reaction *q, double l0, double l) {
  OReaction r{q->type, q->radius, q->E, q->density, q->B, q->length_ratio_trns,
              q->cmp_ratio_trns};
  if (r.type == MF_REACTION_LINEAR) return linear_force(l0, l, r.E, r.area());
  return nonlinear_force(r, l0, l);
}
int oracle_nodal_mass(oracle_batch *b, int rve, double *mass_input_order) {
  ORve &R = b->rves[rve];
  for (int i = 0; i < R.N; ++i) mass_input_order[i] = R.mass[R.perm[i]];
  return 0;
}

}  // extern "C"
