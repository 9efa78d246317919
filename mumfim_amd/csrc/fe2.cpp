// FE^2 message adapter: the micro side of MultiscaleRVEAnalysis::run
// (src/mumfim/microscale/MultiscaleRVEAnalysis.cc:197-305) over the batch
// C ABI.  The macro scale's per-RVE records (MultiscaleMicroFOParams.h:19-120)
// go in; micro_fo_result (sigma, 3 zeros, C) per macro iteration and
// micro_fo_step_result (orientation tensors) per macro step come out.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mf_internal.hpp"

using namespace mf;

namespace {

// field indices of MultiscaleMicroFOParams.h
enum { kRveType = 0 };
enum { kFiberRadius = 0, kVolumeFraction = 1, kAxisX = 5, kAxisY = 6, kAxisZ = 7 };
enum { kViscousDamping = 5 };
enum { kSolverType = 0 };

double total_fiber_length(const mf_network &n) {
  double L = 0.0;
  for (int e = 0; e < n.n_fibers; ++e) {
    const int a = n.fibers[2 * e], b = n.fibers[2 * e + 1];
    const double dx = n.coords[3 * b] - n.coords[3 * a];
    const double dy = n.coords[3 * b + 1] - n.coords[3 * a + 1];
    const double dz = n.coords[3 * b + 2] - n.coords[3 * a + 2];
    L += std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  return L;
}

bool network_ok(const mf_network &n) {
  if (n.n_nodes <= 0 || n.n_fibers < 0 || !n.coords || (n.n_fibers > 0 && !n.fibers))
    return false;
  for (int e = 0; e < 2 * n.n_fibers; ++e)
    if (n.fibers[e] < 0 || n.fibers[e] >= n.n_nodes) return false;
  return true;
}

}  // namespace

struct mf_fo {
  mf_batch *b = nullptr;
  int n = 0;
  std::vector<double> normal;  // 3 per RVE, ORIENTATION_AXIS_{X,Y,Z}
  ~mf_fo() { mf_batch_destroy(b); }
};

extern "C" {

int mf_fo_scale_conversion(const mf_network *net, double fiber_radius, double volume_fraction,
                           double *scale) {
  if (!net || !scale || !network_ok(*net) || !(fiber_radius > 0) || !(volume_fraction > 0)) {
    set_error("mf_fo_scale_conversion: bad arguments");
    return MF_ERR_INVALID;
  }
  // calcRVEDimensionality (MultiscaleCoupling.cc:7-16), pi as 4 atan(1)
  // (MultiscaleRVEAnalysis.cc:162)
  const double pi = 4 * std::atan(1.0);
  const double area = pi * fiber_radius * fiber_radius;
  const double rve_dim = std::sqrt(total_fiber_length(*net) * area / volume_fraction);
  if (!(rve_dim > 0)) {
    set_error("mf_fo_scale_conversion: network has no fiber length");
    return MF_ERR_INVALID;
  }
  *scale = 1.0 / (rve_dim * rve_dim);
  return MF_OK;
}

int mf_fo_create(const mf_network *nets, int32_t n, const int32_t *hdrs, const double *prms,
                 const double *slvr, const int32_t *slvr_int, const mf_solver_params *params,
                 mf_fo **out) {
  if (!out || !nets || n <= 0 || !hdrs || !prms || !slvr || !slvr_int) {
    set_error("mf_fo_create: NULL argument or no RVEs");
    return MF_ERR_INVALID;
  }
  *out = nullptr;
  std::vector<mf_network> desc(nets, nets + n);
  std::vector<double> eta(n);
  for (int r = 0; r < n; ++r) {
    const int32_t *h = hdrs + (size_t)MF_FO_HEADER_FIELDS * r;
    const double *p = prms + (size_t)MF_FO_PARAM_FIELDS * r;
    const double *sv = slvr + (size_t)MF_FO_SOLVER_FIELDS * r;
    const int32_t *si = slvr_int + (size_t)MF_FO_SOLVER_INT_FIELDS * r;
    // MultiscaleRVEAnalysis.cc:144-180: only fiber-only RVEs run batched
    if (h[kRveType] != MF_FO_FIBER_ONLY) {
      set_error(h[kRveType] == MF_FO_ISOTROPIC_NEOHOOKEAN
                    ? "Currently Neohookean Analysis is not supported in batched mode"
                    : "The Microscale/RVE type is not valid");
      return MF_ERR_UNSUPPORTED;
    }
    // the batched analysis reads explicit solution strategies
    // (BatchedFiberRVEAnalysisExplicit.h:213-224)
    if (si[kSolverType] != MF_FO_SOLVER_EXPLICIT) {
      set_error("mf_fo_create: RVE " + std::to_string(r) +
                " does not request the explicit micro solver");
      return MF_ERR_UNSUPPORTED;
    }
    double s = 1.0;
    int e = mf_fo_scale_conversion(&desc[r], p[kFiberRadius], p[kVolumeFraction], &s);
    if (e != MF_OK) {
      set_error("mf_fo_create: RVE " + std::to_string(r) + ": " + mf_last_error());
      return e;
    }
    desc[r].scale_factor = s;
    eta[r] = sv[kViscousDamping];
  }
  mf_solver_params p;
  if (params)
    p = *params;
  else
    mf_default_params(&p);
  mf_fo *f = new mf_fo;
  f->n = n;
  int e = mf_batch_create(desc.data(), n, &p, &f->b);
  if (e == MF_OK) e = mf_batch_set_damping(f->b, eta.data());
  if (e != MF_OK) {
    delete f;
    return e;
  }
  f->normal.resize(3 * (size_t)n);
  for (int r = 0; r < n; ++r) {
    const double *pr = prms + (size_t)MF_FO_PARAM_FIELDS * r;
    f->normal[3 * r] = pr[kAxisX];
    f->normal[3 * r + 1] = pr[kAxisY];
    f->normal[3 * r + 2] = pr[kAxisZ];
  }
  *out = f;
  return MF_OK;
}

void mf_fo_destroy(mf_fo *f) { delete f; }

mf_batch *mf_fo_batch(mf_fo *f) { return f ? f->b : nullptr; }

int mf_fo_run(mf_fo *f, const double *data, double *results, int32_t *status) {
  if (!f || !data || !results) {
    set_error("mf_fo_run: NULL argument");
    return MF_ERR_INVALID;
  }
  const int n = f->n;
  std::vector<double> sigma(6 * (size_t)n), C(36 * (size_t)n);
  std::vector<int32_t> st(n), stc(n);
  // run(F, stress) commits (update_coords defaults to true), then the
  // tangent (MultiscaleRVEAnalysis.cc:257-258)
  int e = mf_batch_run(f->b, data, n, sigma.data(), 1, st.data());
  if (e != MF_OK) return e;
  if ((e = mf_batch_tangent(f->b, C.data(), stc.data())) != MF_OK) return e;
  for (int r = 0; r < n; ++r) {
    double *res = results + (size_t)MF_FO_RESULT * r;
    for (int j = 0; j < 6; ++j) res[j] = sigma[6 * (size_t)r + j];
    for (int j = 6; j < 9; ++j) res[j] = 0.0;  // avgVolStress, unused
    for (int j = 0; j < 36; ++j) res[9 + j] = C[36 * (size_t)r + j];
    if (status) status[r] = st[r] != MF_STATUS_OK ? st[r] : stc[r];
  }
  return MF_OK;
}

int mf_fo_step_results(mf_fo *f, double *step_results) {
  if (!f || !step_results) {
    set_error("mf_fo_step_results: NULL argument");
    return MF_ERR_INVALID;
  }
  const int n = f->n;
  std::vector<double> o3(9 * (size_t)n), o2(9 * (size_t)n);
  // recoverMultiscaleStepResults (MultiscaleCoupling.h:32-52)
  int e = mf_batch_orientation_3d(f->b, o3.data());
  if (e != MF_OK) return e;
  if ((e = mf_batch_orientation_2d(f->b, f->normal.data(), n, o2.data())) != MF_OK) return e;
  for (int r = 0; r < n; ++r) {
    double *s = step_results + (size_t)MF_FO_STEP_RESULT * r;
    for (int j = 0; j < 9; ++j) {
      s[j] = o3[9 * (size_t)r + j];
      s[9 + j] = o2[9 * (size_t)r + j];
    }
  }
  return MF_OK;
}

}  // extern "C"
