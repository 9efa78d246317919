// Checks the packed RVE layout that the FP32c kernels index by: sliced ELL
// padding, the fiber stream, the per-entry fiber references and the
// per-fiber parameters.  Built against the library's packer source
// (mumfim_amd/csrc/network.cpp) by tests/test_pack_layout.py; prints "ok"
// or the first violated property.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>
#include <utility>

#include "../mumfim_amd/csrc/mf_internal.hpp"

namespace mf {
void set_error(const std::string &m) { std::fprintf(stderr, "packer: %s\n", m.c_str()); }
}  // namespace mf

using namespace mf;

static void fail(const char *what, int net) {
  std::printf("FAIL %s (network %d)\n", what, net);
  std::exit(1);
}

static void check(const PackedRVE &p, int t) {
  const int N = p.N, Nf = p.Nf, M = p.M, ne = p.ell_free_entries;
  if ((int)p.fstream.size() != 4 * M || (int)p.fparam.size() != 4 * M) fail("sizes", t);
  if (p.ell_w.size() != p.ell_fref.size() || p.ell_w.size() != p.ell_il.size()) fail("ell sizes", t);
  // ELL of the free slices: real entries name the other end with 1/l0 > 0;
  // padding names the node itself with 1/l0 = 0 and the zero force record
  for (int r = 0; r < Nf; ++r) {
    const int k = r / kSlice;
    for (int j = 0; j < p.slice_D[k]; ++j) {
      const int at = p.slice_base[k] + j * kSlice + r % kSlice;
      if (at < 0 || at >= ne) fail("entry outside the free slices", t);
      const int w = p.ell_w[at] & 0xFFFFFF;
      if (w < 0 || w >= N) fail("entry names no node", t);
      const unsigned fr = (unsigned)p.ell_fref[at];
      if ((fr & 0x7FFFFFFFu) > (unsigned)M) fail("reference range", t);
      if (j >= p.deg[r]) {
        if (w != r || p.ell_il[at] != 0.0f) fail("padding", t);
        if (fr != (unsigned)M) fail("padding reference", t);
      } else if (!(p.ell_il[at] > 0.0f)) {
        fail("1/l0 of a real entry", t);
      }
    }
  }
  // fiber stream: every fiber once; end a is free when either end is; the
  // end slots are the fiber's own ELL entries; fibers with two boundary
  // ends come last; the references point back at the stream position
  std::set<std::pair<int, int>> seen;
  std::set<int> slots;
  bool tail = false;
  for (int e = 0; e < M; ++e) {
    const int *q = &p.fstream[4 * (size_t)e];
    const int a = q[0], b = q[1], sa = q[2], sb = q[3];
    if (a < 0 || a >= N || b < 0 || b >= N || a == b) fail("stream ranks", t);
    if (!seen.insert({std::min(a, b), std::max(a, b)}).second) fail("fiber twice", t);
    if ((a < Nf) != (sa >= 0) || (b < Nf) != (sb >= 0)) fail("slot presence", t);
    if (a >= Nf && b < Nf) fail("end a not the free end", t);
    if (sa < 0) tail = true;
    else if (tail) fail("boundary-only fibers not last", t);
    for (int pass = 0; pass < 2; ++pass) {
      const int r = pass ? b : a, o = pass ? a : b, s = pass ? sb : sa;
      if (s < 0) continue;
      if (!slots.insert(s).second) fail("slot twice", t);
      const int k = r / kSlice, rel = s - p.slice_base[k];
      if (rel < 0 || rel % kSlice != r % kSlice || rel / kSlice >= p.deg[r]) fail("slot row", t);
      if ((p.ell_w[s] & 0xFFFFFF) != o) fail("slot names the other end", t);
      const unsigned fr = (unsigned)p.ell_fref[s];
      if ((fr & 0x7FFFFFFFu) != (unsigned)e || ((fr >> 31) != 0) != (pass == 0))
        fail("reference sign or position", t);
    }
    // parameters: reference vector rounded once from FP64; coefficient by law
    const float *c = &p.fparam[4 * (size_t)e];
    for (int d = 0; d < 3; ++d)
      if (c[d] != (float)(p.X[3 * b + d] - p.X[3 * a + d])) fail("reference vector", t);
    if (sa < 0) {
      if (c[3] != 0.0f) fail("boundary-only coefficient", t);
    } else {
      const Reaction &rq = p.reactions[(uint32_t)p.ell_w[sa] >> 24];
      const double l0 = p.ell_l0[sa];
      const float want = rq.type == MF_REACTION_NONLINEAR
                             ? -(float)(1.0 / l0)
                             : (float)(0.5 * rq.E * rq.area() / (l0 * l0 * l0));
      if (c[3] != want) fail("coefficient", t);
    }
  }
  long free_ends = 0;
  for (int r = 0; r < Nf; ++r) free_ends += p.deg[r];
  if ((long)slots.size() != free_ends) fail("free ends without a slot", t);
  // fibers of both laws: a node lists its nonlinear fibers first; free
  // nodes run by degree, then by nonlinear count (both descending); the
  // stream holds its nonlinear free fibers first
  auto is_nl = [&](int at) {
    return p.reactions[(uint32_t)p.ell_w[at] >> 24].type == MF_REACTION_NONLINEAR;
  };
  int n_nl = 0;
  for (int e = 0; e < M; ++e) n_nl += p.reactions[p.rid[e]].type == MF_REACTION_NONLINEAR;
  if (n_nl == 0 || n_nl == M) return;
  int prev_deg = 1 << 30, prev_nl = 1 << 30;
  for (int r = 0; r < Nf; ++r) {
    const int k = r / kSlice;
    int cnt = 0;
    bool lin_seen = false;
    for (int j = 0; j < p.deg[r]; ++j) {
      const int at = p.slice_base[k] + j * kSlice + r % kSlice;
      if (is_nl(at)) {
        if (lin_seen) fail("nonlinear entry after a linear one", t);
        ++cnt;
      } else {
        lin_seen = true;
      }
    }
    if (p.deg[r] > prev_deg || (p.deg[r] == prev_deg && cnt > prev_nl)) fail("node order", t);
    prev_deg = p.deg[r];
    prev_nl = cnt;
  }
  bool lin_stream = false;
  for (int e = 0; e < M; ++e) {
    const int sa = p.fstream[4 * (size_t)e + 2];
    if (sa < 0) break;
    if (!is_nl(sa)) lin_stream = true;
    else if (lin_stream) fail("nonlinear fiber after a linear one in the stream", t);
  }
}

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 60;
  int checked = 0;
  for (int t = 0; t < n; ++t) {
    mf_gen_params g{};
    g.seed = 1000 + t;
    g.n_nodes = 200 + (t * 137) % 3800;
    g.k_neighbors = 4;
    g.boundary_band = 0.02;
    g.radius = 3.5e-8;
    g.E = 4.32e7;
    g.density = 1000;
    g.nonlinear_fraction = 0.5;
    g.B = 1.2;
    g.length_ratio_trns = 1.05;
    Network net;
    if (generate_network(g, net) != MF_OK) fail("generate", t);
    for (int policy : {MF_REACT_PER_FIBER, MF_REACT_BATCHED_REFERENCE}) {
      PackedRVE p;
      if (pack_rve(net, policy, p) != MF_OK) fail("pack", t);
      check(p, t);
      ++checked;
    }
  }
  std::printf("ok %d\n", checked);
  return 0;
}
