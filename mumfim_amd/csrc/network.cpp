// Host-side network handling: the seeded generator, the reader for the
// reference's text formats and the packer that lays one RVE out in kernel
// order (boundary-last, degree-sorted free nodes, sliced ELL incidences).
#include <algorithm>
#include <cmath>
#include <fstream>
#include <numeric>
#include <sstream>

#include "mf_internal.hpp"

namespace mf {

int network_from_desc(const mf_network &d, Network &out) {
  if (d.n_nodes <= 0 || d.n_fibers < 0 || d.n_boundary < 0 ||
      d.n_boundary > d.n_nodes || d.n_reactions <= 0 ||
      d.n_reactions > kMaxReactions || !d.coords || (!d.fibers && d.n_fibers) ||
      (!d.boundary && d.n_boundary) || !d.reactions) {
    set_error("mf_network: bad counts or NULL arrays");
    return MF_ERR_INVALID;
  }
  out.N = d.n_nodes;
  out.M = d.n_fibers;
  out.Nb = d.n_boundary;
  out.X.assign(d.coords, d.coords + 3 * (size_t)d.n_nodes);
  out.fib.assign(d.fibers, d.fibers + 2 * (size_t)d.n_fibers);
  if (d.fiber_reaction)
    out.rid.assign(d.fiber_reaction, d.fiber_reaction + d.n_fibers);
  else
    out.rid.assign(d.n_fibers, 0);
  out.bnd.assign(d.boundary, d.boundary + d.n_boundary);
  out.reactions.clear();
  for (int i = 0; i < d.n_reactions; ++i) {
    const mf_reaction &r = d.reactions[i];
    Reaction q;
    q.type = r.type;
    q.radius = r.radius;
    q.E = r.E;
    q.rho = r.density;
    q.B = r.B;
    q.lt = r.length_ratio_trns;
    q.lc = r.cmp_ratio_trns;
    if (q.type != MF_REACTION_LINEAR && q.type != MF_REACTION_NONLINEAR) {
      set_error("mf_network: unknown reaction type");
      return MF_ERR_INVALID;
    }
    out.reactions.push_back(q);
  }
  out.scale = d.scale_factor;
  return MF_OK;
}

// ---- generator ---------------------------------------------------------------
// Nodes i.i.d. uniform in [-0.5,0.5)^3; fibers are the undirected, deduplicated
// k-nearest-neighbour edges (ties broken by node index), sorted by (i, j);
// boundary nodes satisfy 0.5 - max_d |X_d| <= band.  Exact k-NN through a
// uniform cell grid searched in growing shells.
namespace {
struct Cand {
  double d2;
  int j;
  bool operator<(const Cand &o) const {
    return d2 < o.d2 || (d2 == o.d2 && j < o.j);
  }
};
}  // namespace

static void knn_edges(const std::vector<double> &X, int N, int k,
                      std::vector<int32_t> &edges) {
  k = std::min(k, N - 1);
  edges.clear();
  if (k <= 0) return;
  int G = std::max(1, (int)std::cbrt(N / 2.0));
  double w = 1.0 / G;
  std::vector<int> cell_of(N), count(G * G * G + 1, 0);
  auto cidx = [&](double v) {
    int c = (int)std::floor((v + 0.5) * G);
    return std::min(std::max(c, 0), G - 1);
  };
  for (int i = 0; i < N; ++i) {
    int cx = cidx(X[3 * i]), cy = cidx(X[3 * i + 1]), cz = cidx(X[3 * i + 2]);
    cell_of[i] = (cz * G + cy) * G + cx;
    count[cell_of[i] + 1]++;
  }
  for (int c = 0; c < G * G * G; ++c) count[c + 1] += count[c];
  std::vector<int> members(N), fill(count.begin(), count.end() - 1);
  for (int i = 0; i < N; ++i) members[fill[cell_of[i]]++] = i;
  std::vector<std::pair<int, int>> pairs;
  pairs.reserve((size_t)N * k);
  std::vector<Cand> best;
  for (int i = 0; i < N; ++i) {
    const double px = X[3 * i], py = X[3 * i + 1], pz = X[3 * i + 2];
    int cx = cidx(px), cy = cidx(py), cz = cidx(pz);
    best.clear();
    for (int r = 0;; ++r) {
      for (int z = cz - r; z <= cz + r; ++z)
        for (int y = cy - r; y <= cy + r; ++y)
          for (int x = cx - r; x <= cx + r; ++x) {
            if (std::max({std::abs(x - cx), std::abs(y - cy), std::abs(z - cz)}) != r)
              continue;
            if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) continue;
            int c = (z * G + y) * G + x;
            for (int m = count[c]; m < count[c + 1]; ++m) {
              int j = members[m];
              if (j == i) continue;
              double dx = X[3 * j] - px, dy = X[3 * j + 1] - py,
                     dz = X[3 * j + 2] - pz;
              Cand cd{dx * dx + dy * dy + dz * dz, j};
              if ((int)best.size() < k) {
                best.push_back(cd);
                std::push_heap(best.begin(), best.end());
              } else if (cd < best.front()) {
                std::pop_heap(best.begin(), best.end());
                best.back() = cd;
                std::push_heap(best.begin(), best.end());
              }
            }
          }
      bool covers = (cx - r <= 0 && cy - r <= 0 && cz - r <= 0 &&
                     cx + r >= G - 1 && cy + r >= G - 1 && cz + r >= G - 1);
      if (covers) break;
      if ((int)best.size() == k) {
        // distance from the point to the outside of the searched block
        double lo[3] = {(cx - r) * w - 0.5, (cy - r) * w - 0.5, (cz - r) * w - 0.5};
        double hi[3] = {(cx + r + 1) * w - 0.5, (cy + r + 1) * w - 0.5,
                        (cz + r + 1) * w - 0.5};
        double p[3] = {px, py, pz};
        double dmin = 1e300;
        for (int d = 0; d < 3; ++d) {
          if (lo[d] > -0.5) dmin = std::min(dmin, p[d] - lo[d]);
          if (hi[d] < 0.5) dmin = std::min(dmin, hi[d] - p[d]);
        }
        if (dmin * dmin >= best.front().d2) break;
      }
    }
    for (const Cand &c : best)
      pairs.emplace_back(std::min(i, c.j), std::max(i, c.j));
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  edges.resize(2 * pairs.size());
  for (size_t e = 0; e < pairs.size(); ++e) {
    edges[2 * e] = pairs[e].first;
    edges[2 * e + 1] = pairs[e].second;
  }
}

int generate_network(const mf_gen_params &g, Network &out) {
  if (g.n_nodes < 2 || g.k_neighbors < 1 || g.boundary_band < 0 ||
      g.radius <= 0 || g.E <= 0 || g.density <= 0) {
    set_error("mf_gen_params: invalid");
    return MF_ERR_INVALID;
  }
  const int N = g.n_nodes;
  out.N = N;
  out.X.resize(3 * (size_t)N);
  SplitMix64 rn(g.seed, kStreamNodes);
  for (int i = 0; i < 3 * N; ++i) out.X[i] = rn.uniform() - 0.5;
  knn_edges(out.X, N, g.k_neighbors, out.fib);
  out.M = (int)(out.fib.size() / 2);
  out.rid.assign(out.M, 0);
  if (g.nonlinear_fraction > 0) {
    SplitMix64 rr(g.seed, kStreamReactions);
    for (int e = 0; e < out.M; ++e)
      out.rid[e] = rr.uniform() < g.nonlinear_fraction ? 1 : 0;
  }
  out.bnd.clear();
  for (int i = 0; i < N; ++i) {
    double m = std::max({std::fabs(out.X[3 * i]), std::fabs(out.X[3 * i + 1]),
                         std::fabs(out.X[3 * i + 2])});
    if (0.5 - m <= g.boundary_band) out.bnd.push_back(i);
  }
  out.Nb = (int)out.bnd.size();
  Reaction lin, nl;
  lin.type = MF_REACTION_LINEAR;
  lin.radius = g.radius;
  lin.E = g.E;
  lin.rho = g.density;
  nl = lin;
  nl.type = MF_REACTION_NONLINEAR;
  nl.B = g.B;
  nl.lt = g.length_ratio_trns;
  nl.lc = 1.0;
  out.reactions = {lin, nl};
  out.scale = 1.0;
  return MF_OK;
}

// ---- reference text formats ------------------------------------------------
// Network file: "nn ne np", nn lines "x y z", ne lines "n0 n1" (0-based),
// np periodic pairs (ignored by the explicit solver), FiberNetworkIO.cc:47-106.
// Params file "<path>.params": "nr ne", nr reaction lines
// ("0 r E rho" linear, "1 r E B lt rho" nonlinear with lc = 1),
// then ne reaction ids, FiberNetworkIO_impl.h:17-39,68-87.
int load_network_file(const std::string &path, Network &out) {
  std::ifstream is(path);
  if (!is) {
    set_error("cannot open " + path);
    return MF_ERR_IO;
  }
  int nn = 0, ne = 0, np = 0;
  if (!(is >> nn >> ne >> np) || nn <= 0 || ne < 0) {
    set_error("bad header in " + path);
    return MF_ERR_IO;
  }
  out.N = nn;
  out.M = ne;
  out.X.resize(3 * (size_t)nn);
  for (int i = 0; i < 3 * nn; ++i)
    if (!(is >> out.X[i])) {
      set_error("truncated coordinates in " + path);
      return MF_ERR_IO;
    }
  out.fib.resize(2 * (size_t)ne);
  for (int i = 0; i < 2 * ne; ++i)
    if (!(is >> out.fib[i]) || out.fib[i] < 0 || out.fib[i] >= nn) {
      set_error("bad fiber line in " + path);
      return MF_ERR_IO;
    }
  out.rid.assign(ne, 0);
  out.reactions.clear();
  std::ifstream ps(path + ".params");
  if (ps) {
    int nr = 0, pe = 0;
    ps >> nr >> pe;
    if (nr <= 0 || nr > kMaxReactions || pe != ne) {
      set_error("params file does not match network " + path);
      return MF_ERR_IO;
    }
    for (int r = 0; r < nr; ++r) {
      Reaction q;
      int tp = -1;
      ps >> tp;
      if (tp == MF_REACTION_LINEAR) {
        ps >> q.radius >> q.E >> q.rho;
      } else if (tp == MF_REACTION_NONLINEAR) {
        ps >> q.radius >> q.E >> q.B >> q.lt >> q.rho;
        q.lc = 1.0;
      } else {
        set_error("unknown reaction type in params of " + path);
        return MF_ERR_IO;
      }
      q.type = tp;
      out.reactions.push_back(q);
    }
    for (int e = 0; e < ne; ++e)
      if (!(ps >> out.rid[e]) || out.rid[e] < 0 || out.rid[e] >= nr) {
        set_error("bad reaction id in params of " + path);
        return MF_ERR_IO;
      }
  } else {
    set_error("missing params file " + path + ".params");
    return MF_ERR_IO;
  }
  // Boundary predicate of RVE(0.5): within 1e-8 of any face (RVE.h:86-95),
  // collected in node order (RVE_impl.h:6-20).
  out.bnd.clear();
  for (int i = 0; i < nn; ++i) {
    bool on = false;
    for (int d = 0; d < 3; ++d)
      on = on || std::fabs(0.5 - out.X[3 * i + d]) < 1e-8 ||
           std::fabs(-0.5 - out.X[3 * i + d]) < 1e-8;
    if (on) out.bnd.push_back(i);
  }
  out.Nb = (int)out.bnd.size();
  out.scale = 1.0;
  return MF_OK;
}

// ---- packer ---------------------------------------------------------------
int pack_rve(const Network &net, int reaction_policy, PackedRVE &p) {
  const int N = net.N, M = net.M, Nb = net.Nb;
  if ((int)net.reactions.size() < 1) {
    set_error("network without reactions");
    return MF_ERR_INVALID;
  }
  std::vector<char> is_bnd(N, 0);
  for (int b : net.bnd) {
    if (b < 0 || b >= N || is_bnd[b]) {
      set_error("boundary list out of range or duplicated");
      return MF_ERR_INVALID;
    }
    is_bnd[b] = 1;
  }
  for (int e = 0; e < M; ++e) {
    int a = net.fib[2 * e], c = net.fib[2 * e + 1];
    if (a < 0 || a >= N || c < 0 || c >= N || a == c) {
      set_error("fiber endpoint out of range or self loop");
      return MF_ERR_INVALID;
    }
    if (net.rid[e] < 0 || net.rid[e] >= (int)net.reactions.size()) {
      set_error("fiber reaction id out of range");
      return MF_ERR_INVALID;
    }
  }
  p.N = N;
  p.M = M;
  p.Nb = Nb;
  p.Nf = N - Nb;
  p.scale = net.scale;
  // reactions after the policy
  if (reaction_policy == MF_REACT_BATCHED_REFERENCE) {
    Reaction r0 = net.reactions[0];
    r0.type = MF_REACTION_LINEAR;
    p.reactions = {r0};
    p.rid.assign(M, 0);
    p.c_sound = std::sqrt(r0.E / r0.rho);
  } else {
    p.reactions = net.reactions;
    p.rid = net.rid;
    double c = 0;
    std::vector<char> used(p.reactions.size(), 0);
    for (int e = 0; e < M; ++e) used[p.rid[e]] = 1;
    for (size_t r = 0; r < p.reactions.size(); ++r)
      if (used[r] || M == 0)
        c = std::max(c, std::sqrt(p.reactions[r].E / p.reactions[r].rho));
    p.c_sound = c;
  }
  // degrees in input order; in an RVE with fibers of both laws, the number of
  // nonlinear incidences of each node
  std::vector<char> fib_nl(M, 0);
  int n_nl_fib = 0;
  for (int e = 0; e < M; ++e) {
    fib_nl[e] = p.reactions[p.rid[e]].type == MF_REACTION_NONLINEAR;
    n_nl_fib += fib_nl[e];
  }
  const bool mixed = n_nl_fib > 0 && n_nl_fib < M;
  std::vector<int> deg_in(N, 0), nl_in(N, 0);
  for (int e = 0; e < M; ++e) {
    deg_in[net.fib[2 * e]]++;
    deg_in[net.fib[2 * e + 1]]++;
    if (mixed && fib_nl[e]) {
      nl_in[net.fib[2 * e]]++;
      nl_in[net.fib[2 * e + 1]]++;
    }
  }
  // kernel order: free nodes (sorted by degree desc, then by nonlinear
  // incidences desc, then Morton order), then boundary nodes in
  // boundary-list order (BatchedReorderMesh.h:54-73).  The second key gives
  // the lanes of a wave similar counts of nonlinear incidences, which the
  // FP32 kernels list first.
  std::vector<int> free_nodes;
  free_nodes.reserve(N - Nb);
  for (int i = 0; i < N; ++i)
    if (!is_bnd[i]) free_nodes.push_back(i);
  // within a degree, Morton order of X: lanes of a wave then gather
  // neighbours that sit close together in LDS
  std::vector<uint32_t> morton(N, 0);
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], net.X[3 * i + d]);
        hi[d] = std::max(hi[d], net.X[3 * i + d]);
      }
    auto spread = [](uint32_t v) {  // 10 bits -> every third of 30 bits
      v &= 0x3FF;
      v = (v | (v << 16)) & 0x030000FF;
      v = (v | (v << 8)) & 0x0300F00F;
      v = (v | (v << 4)) & 0x030C30C3;
      v = (v | (v << 2)) & 0x09249249;
      return v;
    };
    for (int i = 0; i < N; ++i) {
      uint32_t c[3];
      for (int d = 0; d < 3; ++d) {
        const double w = hi[d] > lo[d] ? (net.X[3 * i + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
        c[d] = (uint32_t)std::min(1023.0, std::max(0.0, std::floor(w * 1024.0)));
      }
      morton[i] = spread(c[0]) | (spread(c[1]) << 1) | (spread(c[2]) << 2);
    }
  }
  std::stable_sort(free_nodes.begin(), free_nodes.end(), [&](int a, int b) {
    if (deg_in[a] != deg_in[b]) return deg_in[a] > deg_in[b];
    if (nl_in[a] != nl_in[b]) return nl_in[a] > nl_in[b];
    return morton[a] < morton[b];
  });
  p.input_of_rank.resize(N);
  p.rank_of_input.resize(N);
  for (int r = 0; r < (int)free_nodes.size(); ++r) p.input_of_rank[r] = free_nodes[r];
  for (int k = 0; k < Nb; ++k) p.input_of_rank[p.Nf + k] = net.bnd[k];
  for (int r = 0; r < N; ++r) p.rank_of_input[p.input_of_rank[r]] = r;
  p.X.resize(3 * (size_t)N);
  p.deg.resize(N);
  for (int r = 0; r < N; ++r) {
    int i = p.input_of_rank[r];
    for (int d = 0; d < 3; ++d) p.X[3 * r + d] = net.X[3 * i + d];
    p.deg[r] = deg_in[i];
  }
  // fibers: ranks, reference length (getElementLength, ...Base.h:203-215)
  p.fib.resize(2 * (size_t)M);
  p.l0.resize(M);
  double sum_l0 = 0;
  int n_nl = 0;
  for (int e = 0; e < M; ++e) {
    int a = net.fib[2 * e], c = net.fib[2 * e + 1];
    p.fib[2 * e] = p.rank_of_input[a];
    p.fib[2 * e + 1] = p.rank_of_input[c];
    double x1 = net.X[3 * c] - net.X[3 * a];
    double x2 = net.X[3 * c + 1] - net.X[3 * a + 1];
    double x3 = net.X[3 * c + 2] - net.X[3 * a + 2];
    p.l0[e] = std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    sum_l0 += p.l0[e];
    if (p.reactions[p.rid[e]].type == MF_REACTION_NONLINEAR) n_nl++;
  }
  p.mean_l0 = M ? sum_l0 / M : 0.0;
  p.nonlinear_frac = M ? (double)n_nl / M : 0.0;
  // lumped mass m_v = sum over incident fibers of rho*A*l0/2, accumulated in
  // fiber order (MassIntegrator.cc:16-55 with linear shape functions)
  p.mass.assign(N, 0.0);
  for (int e = 0; e < M; ++e) {
    const Reaction &r = p.reactions[p.rid[e]];
    double c = r.rho * r.area() * p.l0[e] * 0.5;
    p.mass[p.fib[2 * e]] += c;
    p.mass[p.fib[2 * e + 1]] += c;
  }
  p.max_free_deg = 0;
  p.n_split = 0;
  for (int r = 0; r < p.Nf; ++r) {
    p.max_free_deg = std::max(p.max_free_deg, p.deg[r]);
    if (p.deg[r] > kLaneDegree) p.n_split++;
  }
  // CSR in kernel order, entries in increasing fiber index; with fibers of
  // both laws, a node's nonlinear fibers first (each law in increasing fiber
  // index)
  std::vector<int> row(N + 1, 0);
  for (int r = 0; r < N; ++r) row[r + 1] = row[r] + p.deg[r];
  std::vector<int> fillp(row.begin(), row.end() - 1);
  std::vector<int32_t> csr_w(row[N]), csr_e(row[N]);
  std::vector<double> csr_l0(row[N]);
  std::vector<int> fib_order;
  fib_order.reserve(M);
  for (int pass = mixed ? 0 : 1; pass < 2; ++pass)
    for (int e = 0; e < M; ++e)
      if (!mixed || (fib_nl[e] != 0) == (pass == 0)) fib_order.push_back(e);
  for (int e : fib_order) {
    int a = p.fib[2 * e], c = p.fib[2 * e + 1];
    int32_t tag = p.rid[e] << 24;
    csr_e[fillp[a]] = e;
    csr_w[fillp[a]] = c | tag;
    csr_l0[fillp[a]++] = p.l0[e];
    csr_e[fillp[c]] = e;
    csr_w[fillp[c]] = a | tag;
    csr_l0[fillp[c]++] = p.l0[e];
  }
  // sliced ELL
  const int n_slices = (N + kSlice - 1) / kSlice;
  p.slice_base.resize(n_slices);
  p.slice_D.resize(n_slices);
  int total = 0;
  const int n_free_slices = (p.Nf + kSlice - 1) / kSlice;
  for (int k = 0; k < n_slices; ++k) {
    int D = 0;
    for (int r = k * kSlice; r < std::min(N, (k + 1) * kSlice); ++r)
      D = std::max(D, p.deg[r]);
    p.slice_base[k] = total;
    p.slice_D[k] = D;
    total += D * kSlice;
    if (k + 1 == n_free_slices) p.ell_free_entries = total;
  }
  if (n_free_slices == 0) p.ell_free_entries = 0;
  p.ell_w.assign(total, 0);
  p.ell_l0.assign(total, 1.0);
  p.ell_il.assign(total, 0.0f);
  for (int k = 0; k < n_slices; ++k)  // padding: the node itself
    for (int j = 0; j < p.slice_D[k]; ++j)
      for (int lane = 0; lane < kSlice; ++lane)
        p.ell_w[(size_t)p.slice_base[k] + (size_t)j * kSlice + lane] =
            k * kSlice + lane < N ? k * kSlice + lane : 0;
  for (int r = 0; r < N; ++r) {
    int k = r / kSlice, lane = r % kSlice;
    for (int j = 0; j < p.deg[r]; ++j) {
      size_t at = (size_t)p.slice_base[k] + (size_t)j * kSlice + lane;
      p.ell_w[at] = csr_w[row[r] + j];
      p.ell_l0[at] = csr_l0[row[r] + j];
      p.ell_il[at] = (float)(1.0 / csr_l0[row[r] + j]);
    }
  }
  // fiber stream: ELL entry of every free fiber end, then fibers in entry
  // order at their first free end, then boundary-only fibers
  std::vector<int32_t> end_slot(2 * (size_t)M, -1);
  for (int r = 0; r < p.Nf; ++r) {
    int k = r / kSlice, lane = r % kSlice;
    for (int j = 0; j < p.deg[r]; ++j) {
      const int e = csr_e[row[r] + j];
      end_slot[2 * (size_t)e + (p.fib[2 * e] == r ? 0 : 1)] =
          p.slice_base[k] + j * kSlice + lane;
    }
  }
  p.fstream.clear();
  p.fstream.reserve(4 * (size_t)M);
  std::vector<char> taken(M, 0);
  for (int k = 0; k < n_free_slices; ++k)
    for (int j = 0; j < p.slice_D[k]; ++j)
      for (int lane = 0; lane < kSlice; ++lane) {
        const int r = k * kSlice + lane;
        if (r >= p.Nf || j >= p.deg[r]) continue;
        const int e = csr_e[row[r] + j];
        if (taken[e]) continue;
        taken[e] = 1;
        const int end = p.fib[2 * e] == r ? 0 : 1;
        const int o = p.fib[2 * e + 1 - end];
        p.fstream.insert(p.fstream.end(),
                         {r, o, end_slot[2 * (size_t)e + end], end_slot[2 * (size_t)e + 1 - end]});
      }
  if (mixed) {
    // nonlinear fibers first (each law in the order above): a wave of the
    // fiber-parallel kernels then holds fibers of one law, except the one
    // wave at the border
    const size_t nfree = p.fstream.size() / 4;
    std::vector<int32_t> part;
    part.reserve(p.fstream.size());
    for (int pass = 0; pass < 2; ++pass)
      for (size_t i = 0; i < nfree; ++i) {
        const int32_t *q = &p.fstream[4 * i];
        const bool nl = p.reactions[(uint32_t)p.ell_w[q[2]] >> 24].type == MF_REACTION_NONLINEAR;
        if (nl == (pass == 0)) part.insert(part.end(), q, q + 4);
      }
    p.fstream.swap(part);
  }
  for (int e = 0; e < M; ++e)
    if (!taken[e]) p.fstream.insert(p.fstream.end(), {p.fib[2 * e], p.fib[2 * e + 1], -1, -1});
  p.fparam.assign(4 * (size_t)M, 0.0f);
  for (int e = 0; e < M; ++e) {
    const int ra = p.fstream[4 * (size_t)e], rb = p.fstream[4 * (size_t)e + 1];
    const int sa = p.fstream[4 * (size_t)e + 2];
    float *q = &p.fparam[4 * (size_t)e];
    for (int d = 0; d < 3; ++d) q[d] = (float)(p.X[3 * rb + d] - p.X[3 * ra + d]);
    if (sa >= 0) {
      const Reaction &rq = p.reactions[(uint32_t)p.ell_w[sa] >> 24];
      const double l0 = p.ell_l0[sa];
      q[3] = rq.type == MF_REACTION_NONLINEAR ? -p.ell_il[sa]
                                              : (float)(0.5 * rq.E * rq.area() / (l0 * l0 * l0));
    }
  }
  p.ell_fref.assign(total, M);
  for (int e = 0; e < M; ++e) {
    const int sa = p.fstream[4 * (size_t)e + 2], sb = p.fstream[4 * (size_t)e + 3];
    if (sa >= 0) p.ell_fref[sa] = (int32_t)((uint32_t)e | 0x80000000u);
    if (sb >= 0) p.ell_fref[sb] = e;
  }
  return MF_OK;
}

}  // namespace mf
