// Internal declarations shared by the host packer, the C ABI and the HIP
// kernels.  Nothing here is part of the public interface.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mumfim_amd.h"

namespace mf {

// Error reporting for the C ABI (thread-local last error).
void set_error(const std::string &msg);

// ---- SplitMix64 streams ---------------------------------------------------
// Each (seed, stream) pair is an independent sequence; doubles carry the top
// 53 bits so the values are the same on every platform.
struct SplitMix64 {
  uint64_t s;
  static uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  SplitMix64(uint64_t seed, uint64_t stream) : s(mix(seed ^ mix(stream))) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

// Stream ids used by the generator.
enum : uint64_t { kStreamNodes = 1, kStreamReactions = 2 };

struct Reaction {
  int type = MF_REACTION_LINEAR;
  double radius = 0, E = 0, rho = 0, B = 0, lt = 0, lc = 1.0;
  // A = r*r*pi, evaluated as in FiberNetworkIO_impl.h:26,36.
  double area() const { return radius * radius * 3.14159265358979323846; }
};

// Owned copy of one network in input order.
struct Network {
  int N = 0, M = 0, Nb = 0;
  std::vector<double> X;        // 3N
  std::vector<int32_t> fib;     // 2M
  std::vector<int32_t> rid;     // M
  std::vector<int32_t> bnd;     // Nb
  std::vector<Reaction> reactions;
  double scale = 1.0;
};

int network_from_desc(const mf_network &d, Network &out);
int generate_network(const mf_gen_params &g, Network &out);
int load_network_file(const std::string &path, Network &out);

// Largest reaction table and ELL degree the kernels accept.
constexpr int kMaxReactions = 8;
constexpr int kSlice = 64;  // nodes per ELL slice = one wavefront
// Fibers per lane in the register-resident FP32 kernel; nodes of higher
// degree (up to twice this) take two lanes.
constexpr int kLaneDegree = 6;

// One RVE packed in kernel order: free nodes sorted by degree (descending)
// and within a degree by the Morton code of X, then boundary nodes in
// boundary-list order.
struct PackedRVE {
  int N = 0, Nf = 0, Nb = 0, M = 0;
  std::vector<int32_t> rank_of_input;  // input node -> kernel rank
  std::vector<int32_t> input_of_rank;
  std::vector<double> X;               // 3N, kernel order
  std::vector<double> mass;            // N
  std::vector<int32_t> deg;            // N
  // Sliced ELL incidence lists: slice k holds ranks [64k, 64k+64); entry j of
  // rank r sits at slice_base[k] + 64*j + (r & 63).  Entries of one node are
  // its incident fibers in increasing fiber index (the reference scatter
  // order).  Packed word = other-end rank | reaction id << 24.  Padding
  // entries (j >= degree) name the node itself with l0 = 1 and 1/l0 = 0, so a
  // kernel that evaluates them gets a zero force.
  std::vector<int32_t> slice_base, slice_D;
  std::vector<int32_t> ell_w;
  std::vector<double> ell_l0;
  std::vector<float> ell_il;           // FP32 1/l0 (0 at padding)
  int ell_free_entries = 0;            // entries of slices that hold free nodes
  int max_free_deg = 0;                // largest degree of a free node
  int n_split = 0;                     // free nodes with degree > kLaneDegree
  // Fiber stream of the fiber-parallel kernel, 4 per fiber: ranks (a, b)
  // and the ELL entries of the fiber in a's and b's lists (-1 at a boundary
  // end).  Fibers are taken in ELL entry order at their first free end, so
  // consecutive fibers mostly have lane-consecutive a; fibers with two
  // boundary ends come last.
  std::vector<int32_t> fstream;
  // Per ELL entry: the entry's fiber as its fiber-stream position (the index
  // of its force record in the fiber-streamed kernel and of its t in the FP64
  // fiber-once kernel), bit 31 set where the node is the fiber's end a (it
  // receives -force).  Padding and boundary-node entries name the zero
  // record at position M.
  std::vector<int32_t> ell_fref;
  // Per fiber in stream order, 4 floats: the reference vector X_b - X_a
  // rounded once to FP32, and the law coefficient (linear: E A / (2 l0^3);
  // nonlinear: -1/l0; 0 for fibers with two boundary ends).
  std::vector<float> fparam;
  std::vector<int32_t> fib;            // 2M kernel ranks, input fiber order
  std::vector<double> l0;              // M
  std::vector<int32_t> rid;            // M (after the reaction policy)
  std::vector<Reaction> reactions;     // after the reaction policy
  double scale = 1.0;
  double c_sound = 0.0;                // sqrt(E/rho) used for the time step
  // Mean fiber length, fiber density, nonlinear fraction (surrogate inputs).
  double mean_l0 = 0.0, nonlinear_frac = 0.0;
};

int pack_rve(const Network &net, int reaction_policy, PackedRVE &out);

// ---- device-side layout ------------------------------------------------------
struct RveDev {
  int32_t N, Nf, Nb, M;
  int64_t node_off;   // first node of this RVE in the node arrays
  int64_t ell_off;    // first ELL entry of this RVE
  int64_t fib_off;    // first fiber of this RVE
  int32_t slice_off;  // first slice of this RVE in the slice tables
  int32_t n_slices;
  int32_t ell_free_entries;
  int32_t react_off;
  int32_t n_react;
  int32_t m_free;     // fibers with a free end (the fiber-stream prefix)
  double c_sound, scale;
  // Reference coordinates in fixed point for the FP32 kernel: X_q =
  // rint(X * xq_scale) fits int32; xq_scale is a power of two.
  double xq_scale;
  float xq_inv;
  // Free nodes of degree above the per-lane fiber count of the
  // register-resident kernel; they are the first n_split free ranks.
  int32_t n_split;
};

// Reaction constants as the kernels consume them.
struct ReactDev {
  int32_t type;
  int32_t pad_;
  // FP64 reference-formula inputs.
  double E, A, B, lt, lc;
  // FP32 algebraic forms: t = f/l as a function of e2 = lambda^2 - 1.
  float EA_half;   // linear: t = EA/2 * il * e2
  float alpha;     // nonlinear alpha = E*A/B
  float B_half;    // B/2
  float lt2m1;     // lambda_t^2 - 1
  float lc2m1;     // lambda_c^2 - 1
  // extensions f = K (lam - L0) + b beyond lambda_t and lambda_c, as
  // t = il (K - C / lam) with C = K L0 - b
  float Kt, Ct;
  float Kc, Cc;
  // alpha expm1(B/2 y) = y P(y) in the toe range: P's coefficients
  // alpha (B/2)^k / k!, k = 1..7 (the Taylor polynomial with both scalings
  // folded in)
  float pe[7];
};

}  // namespace mf
