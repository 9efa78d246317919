// FP32 + compensated explicit solve: FP32 stretch, force and update
// arithmetic, (hi, lo) FP32 positions advanced with TwoSum, FP64 stress.
#include <type_traits>

#include "kernels_common.hpp"

namespace mf {
namespace kern {

// Reaction modes: RM_TABLE looks each fiber's law up in the packed table;
// RM_LIN: the RVE has one linear reaction; RM_MIX2: at most one linear and
// one nonlinear reaction, toe range within the expm1 polynomial's bound.
enum { RM_TABLE = 0, RM_LIN = 1, RM_MIX2 = 2 };

// Pointers into LDS (address space 3): byte addresses within the
// workgroup's LDS allocation.
typedef __attribute__((address_space(3))) const unsigned char lds_uchar;
// Neighbour records are read whole (16 bytes): on gfx950 a 16-byte LDS
// gather runs 13-18% faster in the register kernel than the 12-byte gather of
// the same x, y, z.
typedef float v4f __attribute__((ext_vector_type(4), aligned(16)));
typedef __attribute__((address_space(3))) const v4f lds_v4f;
typedef int v4i __attribute__((ext_vector_type(4), aligned(16)));

// ============================================================================
// FP32 + compensated kernel
// ============================================================================
// State per free node: displacement u = (hi, lo) in FP32 advanced with
// Fast2Sum, velocity in FP32.  LDS holds the displacement hi
// parts (double-buffered) and the reference coordinates X (FP32).  A fiber
// seen from node v towards w has reference vector d0 = X_w - X_v,
// displacement difference delta = u_w - u_v, current vector d = d0 + delta
// and l^2 - l0^2 = (d0 + d) . delta, which keeps full relative precision at
// small strain (the direct form d.d - l0^2 loses eps/strain).
// The reaction mode is read per RVE (linear-only, one linear and one
// nonlinear reaction, or the packed table) and is uniform over the
// workgroup.
template <int B, int NPT, int FL>
__global__ __launch_bounds__(B) void solve32c(SolveArgs a) {
  constexpr bool kEll = (FL & kEllInLds) != 0;
  constexpr bool kDbuf = (FL & kDoubleBuf) != 0;
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int64_t nb = R.node_off;

  size_t off = 0;
  float4 *pos = reinterpret_cast<float4 *>(smem);  // [1 or 2][Np] displacement hi
  off += sizeof(float4) * (kDbuf ? 2 : 1) * Np;
  int4 *Xs = reinterpret_cast<int4 *>(smem + off);  // [Np] reference X, fixed point
  off += sizeof(int4) * Np;
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);  // [Nbp][6] u_b0, g
  off += sizeof(double) * 6 * Nbp;
  float4 *rtab = reinterpret_cast<float4 *>(smem + off);  // packed reaction table
  off += sizeof(float4) * kReactTabStride * kMaxReactions;
  float *ils = reinterpret_cast<float *>(smem + off);
  off += kEll ? sizeof(float) * R.ell_free_entries : 0;
  int32_t *ws = reinterpret_cast<int32_t *>(smem + off);

  const int32_t *gw = a.ell_w + R.ell_off;
  const float *gil = a.ell_il + R.ell_off;
  if (kEll) {
    for (int i = tid; i < R.ell_free_entries; i += B) {
      ws[i] = gw[i];
      ils[i] = gil[i];
    }
  }
  const int32_t *Ew = kEll ? ws : gw;
  const float *Eil = kEll ? ils : gil;
  // HBM lists as raw buffers of the RVE's free-node entries (loads past the
  // end return 0; every load here stays within the slices)
  const int list_bytes = R.ell_free_entries * (int)sizeof(int32_t);
  const __amdgpu_buffer_rsrc_t rsw =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t *>(gw), 0, list_bytes, kBufRsrcWord3);
  const __amdgpu_buffer_rsrc_t rsil =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(gil), 0, list_bytes, kBufRsrcWord3);
  if (tid < R.n_react) pack_react_table(a.reacts[R.react_off + tid], rtab + kReactTabStride * tid);
  int nl_rid = -1;
  float ea_half_lin = 0.0f;
  const bool mix2 = mix2_table(a.reacts + R.react_off, R.n_react, nl_rid, ea_half_lin);
  const ReactDev nlq = a.reacts[R.react_off + (nl_rid < 0 ? 0 : nl_rid)];
  const bool lin_only =
      R.n_react == 1 && a.reacts[R.react_off].type == MF_REACTION_LINEAR;
  const float lin_EA_half = a.reacts[R.react_off].EA_half;

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;
  const double inv_S = 1.0 / a.n_steps;

  float uh[NPT][3], ul[NPT][3], vr[NPT][3], imr[NPT];
  // slice width and base are wave-uniform (scalar registers); Dn: with one
  // linear and one nonlinear reaction, the entries [0, Dn) hold the wave's
  // nonlinear fibers (nodes list them first), later entries are linear on
  // every lane
  int Dr[NPT], sbase[NPT], Dn[NPT];
  // initial displacement: relax adds the affine increment (F - I) x to u,
  // ramp keeps u and sets the boundary increment aside.
  auto init_node = [&](int r, double (&u)[3]) {
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
    for (int k = 0; k < 3; ++k) u[k] = ramp ? up[k] : inc[k] + up[k];
    const float4 p = make_float4((float)u[0], (float)u[1], (float)u[2], 0.0f);
    pos[r] = p;
    if (kDbuf) pos[Np + r] = p;
    Xs[r] = make_int4((int)rint(a.X[3 * g] * R.xq_scale),
                      (int)rint(a.X[3 * g + 1] * R.xq_scale),
                      (int)rint(a.X[3 * g + 2] * R.xq_scale), 0);
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (!ramp) inc[k] = 0.0;
      bnd_set(bst + 6 * b, u, inc);
    }
  };
  const float eta = (float)(a.eta_rve ? a.eta_rve[rv] : a.eta);
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    double u[3] = {0, 0, 0};
    if (r < N) init_node(r, u);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      uh[s][d] = (float)u[d];
      ul[s][d] = (float)(u[d] - (double)uh[s][d]);
      vr[s][d] = 0.0f;
    }
    imr[s] = 0.0f;
    if (r < Nf) imr[s] = (float)(1.0 / a.mass[nb + r]);
    // lane 0 of the wave names the slice; waves past Nf never gather
    const int k = r / kWave;
    Dr[s] = __builtin_amdgcn_readfirstlane(r < Nf ? a.slice_D[R.slice_off + k] : 0);
    sbase[s] = __builtin_amdgcn_readfirstlane(r < Nf ? a.slice_base[R.slice_off + k] : 0);
    int cnt = 0;
    if (mix2 && r < Nf) {
      const int dg = a.deg[nb + r];
      while (cnt < dg && (int)((uint32_t)gw[sbase[s] + lane + cnt * kWave] >> 24) == nl_rid) ++cnt;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt = max(cnt, __shfl_xor(cnt, o, kWave));
    Dn[s] = __builtin_amdgcn_readfirstlane(cnt);
  }
  for (int r = NPT * B + tid; r < N; r += B) {
    double u[3];
    init_node(r, u);
  }
  __syncthreads();

  const float inv_c = (float)(1.0 / R.c_sound);
  const float xq_inv = R.xq_inv;
  int cur = 0;
  // internal force of a slot's node.  ELL padding names the node itself
  // with 1/l0 = 0 (zero force), so every entry of the slice width is
  // evaluated unmasked; the list words and 1/l0 of entries j+1 and j+2 load
  // while entry j is evaluated (rolling prefetch: the lists stream from HBM
  // for large RVEs).  Both ends enter as their FP32-rounded displacements,
  // like the register and fiber kernels (the low parts only keep the
  // accumulation exact).
  auto node_force = [&](int s, int buf, float (&f)[3]) {
    const float4 *P = pos + (kDbuf ? buf * Np : 0);
    const int4 Xv = Xs[s * B + tid];  // own reference coordinate
    const int base = sbase[s] + lane;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int D = Dr[s];
    if (D > 0) {
      // entry jj of the lane's list: from LDS, or from HBM through the
      // buffer resources, whose lane offset is fixed for the slot and whose
      // entry offset is a scalar (no per-entry address arithmetic)
      const int vo = base * (int)sizeof(int32_t);
      auto entry = [&](int jj, int32_t &wv, float &il) __attribute__((always_inline)) {
        if constexpr (kEll) {
          wv = Ew[base + jj * kWave];
          il = Eil[base + jj * kWave];
        } else {
          const int so = jj * (kWave * (int)sizeof(int32_t));
          wv = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rsw, vo, so, 0);
          il = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsil, vo, so, 0));
        }
      };
      // the prefetch index is clamped to the slice's last entry (entries
      // past it reload that one), so the loads need no branch
      const int Dm = D - 1;
      int32_t wv_n, wv_nn;
      float il_n, il_nn;
      entry(0, wv_n, il_n);
      entry(min(1, Dm), wv_nn, il_nn);
      // entries [j0, j1) under one law, chosen outside the loop: the loop
      // body then carries no mode tests and no divergent branch
      auto run = [&](int j0, int j1, auto LAWC) __attribute__((always_inline)) {
        constexpr int LAW = decltype(LAWC)::value;
        for (int j = j0; j < j1; ++j) {
          const int32_t wv = wv_n;
          const float il = il_n;
          wv_n = wv_nn;
          il_n = il_nn;
          entry(min(j + 2, Dm), wv_nn, il_nn);
          const int w = wv & 0xFFFFFF;
          const float4 q = ld16(P + w);
          const int4 xq = ld16(Xs + w);
          // reference vector: exact integer difference, one rounding to FP32
          const float ex = (float)(xq.x - Xv.x) * xq_inv;
          const float ey = (float)(xq.y - Xv.y) * xq_inv;
          const float ez = (float)(xq.z - Xv.z) * xq_inv;
          const float gx = q.x - uh[s][0];
          const float gy = q.y - uh[s][1];
          const float gz = q.z - uh[s][2];
          const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
          // l^2 - l0^2 = (d0 + d) . delta
          const float sl = dot3f(ex + dx, ey + dy, ez + dz, gx, gy, gz);
          const float e2 = sl * (il * il);
          float t;
          if (LAW == kLawLin) {
            t = lin_EA_half * il * e2;
          } else if (LAW == kLawMixNl) {
            // both laws on every lane, then a select
            const float tl = ea_half_lin * il * e2;
            const float tn = toe_ext_t(nlq, e2, il);
            t = (int)((uint32_t)wv >> 24) == nl_rid ? tn : tl;
          } else if (LAW == kLawMixLin) {
            t = ea_half_lin * il * e2;
          } else {
            t = react32_tab(rtab, (uint32_t)wv >> 24, e2, il);
          }
          fx -= dx * t;
          fy -= dy * t;
          fz -= dz * t;
        }
      };
      if (lin_only) {
        run(0, D, std::integral_constant<int, kLawLin>{});
      } else if (mix2) {
        const int dn = min(Dn[s], D);
        run(0, dn, std::integral_constant<int, kLawMixNl>{});
        run(dn, D, std::integral_constant<int, kLawMixLin>{});
      } else {
        run(0, D, std::integral_constant<int, kLawTab>{});
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  auto min_length2 = [&](int buf) {
    const float4 *P = pos + (kDbuf ? buf * Np : 0);
    float lmin = 3.0e38f;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += B) {
      const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
      const float4 p1 = P[n1], p2 = P[n2];
      const int4 X1 = Xs[n1], X2 = Xs[n2];
      const float x1 = (float)(X2.x - X1.x) * xq_inv + (p2.x - p1.x);
      const float x2 = (float)(X2.y - X1.y) * xq_inv + (p2.y - p1.y);
      const float x3 = (float)(X2.z - X1.z) * xq_inv + (p2.z - p1.z);
      lmin = fminf(lmin, dot3f(x1, x2, x3, x1, x2, x3));
    }
    return (double)lmin;
  };

  // The closing half kick of step k and the opening half kick of step k+1
  // use the same acceleration: v_{k+1/2} = v_{k-1/2} + (h_k + h_{k+1}) a_k,
  // applied as soon as a_k and h_{k+1} are known (v_{-1/2} = 0, h_{-1} = 0),
  // so no acceleration is carried across the barrier.
  const float crit_half = (float)(a.crit * 0.5);
  float h_prev = 0.0f;
  float fint[NPT][3];
  {
    double resid_local = 0.0;
#pragma unroll
    for (int s = 0; s < NPT; ++s)
      if (s * B + tid < Nf) node_force(s, cur, fint[s]);
    double lmin2 = min_length2(cur);
    block_min_sum<B>(lmin2, resid_local, red);
    const float h = crit_half * (sqrtf((float)lmin2) * inv_c);
#pragma unroll
    for (int s = 0; s < NPT; ++s)
      if (s * B + tid < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[s][d] = vr[s][d] + h * (-fint[s][d] * imr[s] - eta * vr[s][d]);
      }
    h_prev = h;
  }

  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  for (;;) {
    const float h2 = 2.0f * h_prev;
    const int nxt = kDbuf ? cur ^ 1 : 0;
    float4 *Pn = pos + nxt * Np;
    if (!kDbuf) __syncthreads();  // all force reads of the last step done
    // drift: u (hi, lo) += 2h v with Fast2Sum, exact when |u_hi| >= |increment|;
    // otherwise u is within a few increments of zero and the uncompensated
    // rounding is below eps * |increment|.
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
        float pc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float y = ul[s][d] + h2 * vr[s][d];
          const float sh = uh[s][d] + y;
          ul[s][d] = y - (sh - uh[s][d]);
          uh[s][d] = sh;
          pc[d] = sh;
        }
        Pn[r] = make_float4(pc[0], pc[1], pc[2], 0.0f);
      }
    }
    if (ramp && tid < Nb) {  // waves without boundary nodes skip the amplitude
      const double alpha = ramp_alpha_step(step + 1, inv_S, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
        float ub[3];
        bnd_step(bst + 6 * b, alpha, ub);
        Pn[Nf + b] = make_float4(ub[0], ub[1], ub[2], 0.0f);
      }
    }
    __syncthreads();
    cur = nxt;
    const bool refresh = (step % freq) == 0;
    // the step size follows the new positions and is known before the
    // forces, so each node kicks as soon as its force is done
    float h = h_prev;
    if (refresh) {
      double lm = min_length2(cur), none = 0.0;
      block_min_sum<B>(lm, none, red);
      h = crit_half * (sqrtf((float)lm) * inv_c);
    }
    const float hk = h_prev + h;
    double rsum = 0.0;
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
        float f[3];
        node_force(s, cur, f);
        if (refresh) rsum += (double)dot3f(f[0], f[1], f[2], f[0], f[1], f[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[s][d] = vr[s][d] + hk * (-f[d] * imr[s] - eta * vr[s][d]);
      }
    }
    if (refresh) {
      // slots 2-3 of the scratch: waves may still read slots 0-1 above
      double none = 0.0;
      block_min_sum<B>(none, rsum, red + 2);
      residual = sqrt(rsum);
    }
    h_prev = h;
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

  // finalize: with two buffers the spare one publishes the low parts so the
  // boundary gathers see full displacements; with one, the hi parts only.
  float4 *Ph = pos + (kDbuf ? cur * Np : 0);
  float4 *Pl = pos + (kDbuf ? (cur ^ 1) * Np : 0);
  __syncthreads();  // every read of the other buffer and of red is done
  if (kDbuf) {
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      const int r = s * B + tid;
      if (r < Nf) Pl[r] = make_float4(ul[s][0], ul[s][1], ul[s][2], 0.0f);
    }
  }
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += B) {
    double ub[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    const float4 hb = make_float4((float)ub[0], (float)ub[1], (float)ub[2], 0.0f);
    Ph[Nf + b] = hb;
    if (kDbuf)
      Pl[Nf + b] = make_float4((float)(ub[0] - (double)hb.x), (float)(ub[1] - (double)hb.y),
                               (float)(ub[2] - (double)hb.z), 0.0f);
  }
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    const int64_t gr = nb + r;
    double ub[3], xb[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    for (int d = 0; d < 3; ++d) xb[d] = a.X[3 * gr + d] + ub[d];
    const float4 hb = Ph[r];
    const int4 Xb = Xs[r];
    const float4 lb = kDbuf ? Pl[r] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int k = r / kWave;
    const int base = a.slice_base[R.slice_off + k] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const float il = gil[idx];
      const int w = wv & 0xFFFFFF;
      const float4 qh = Ph[w];
      const int4 xq = Xs[w];
      const float4 ql = kDbuf ? Pl[w] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const float ex = (float)(xq.x - Xb.x) * xq_inv;
      const float ey = (float)(xq.y - Xb.y) * xq_inv;
      const float ez = (float)(xq.z - Xb.z) * xq_inv;
      const float gx = (qh.x - hb.x) + (ql.x - lb.x);
      const float gy = (qh.y - hb.y) + (ql.y - lb.y);
      const float gz = (qh.z - hb.z) + (ql.z - lb.z);
      const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
      const float sl = dot3f(ex + dx, ey + dy, ez + dz, gx, gy, gz);
      const float e2 = sl * (il * il);
      const float t = lin_only ? lin_EA_half * il * e2
                               : react32_tab(rtab, (uint32_t)wv >> 24, e2, il);
      f1 -= dx * t;
      f2 -= dy * t;
      f3 -= dz * t;
    }
    const double g1 = f1, g2 = f2, g3 = f3;
    st[0] += xb[0] * g1;
    st[1] += xb[1] * g2;
    st[2] += xb[2] * g3;
    st[3] += 0.5 * fma(xb[1], g3, xb[2] * g2);
    st[4] += 0.5 * fma(xb[0], g3, xb[2] * g1);
    st[5] += 0.5 * fma(xb[0], g2, xb[1] * g1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * gr + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * gr + d] = ub[d];
  }
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      const int64_t g = nb + r;
      for (int d = 0; d < 3; ++d) {
        const double u = (double)uh[s][d] + (double)ul[s][d];
        if (flags & kWriteX) a.x_out[3 * g + d] = a.X[3 * g + d] + u;
        if (flags & kWriteU) a.u_out[3 * g + d] = u;
      }
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
// FP32 + compensated kernel with per-lane fiber lists in registers.
// Free nodes (degree-sorted, descending) map to lanes, NPT slots per lane:
// slot 0 covers ranks [0, B - S), slot s >= 1 ranks [sB - S, (s+1)B - S),
// where S nodes of degree above kLaneDeg (the first ranks) take two adjacent
// lanes of slot 0 (even/odd) that each gather half of the fibers and exchange
// the partial forces with one DPP quad swap.  Every slot therefore holds at
// most kLaneDeg fibers: the LDS byte offset of the other end, the exactly
// rounded FP32 reference vector d0 = X_w - X_v and either EA/(2 l0^3)
// (linear-only RVEs, t = that * s) or 1/l0.  Entries past a slot's count
// address the node's own row with zero constants and contribute exactly
// zero, so each wave runs its uniform trip count without masks.  A step
// reads one neighbour row (FP32 displacement hi) per fiber from LDS.
// ============================================================================
constexpr int kLaneDeg = kLaneDegree;

// Reaction modes (RM_* above) of the register kernel: RM_LIN keeps
// EA/(2 l0^3) per fiber; RM_MIX2 keeps EA/(2 l0^3) for linear fibers and
// -1/l0 for nonlinear ones, with the nonlinear law's constants uniform.

// Waves per SIMD the register kernel is compiled for (minimum, maximum):
// one node per thread at 256 or 512 threads in the linear and
// one-linear-one-nonlinear modes runs four (two or four solves per CU, at
// most 128 VGPRs); the reaction-table mode and several nodes per thread run
// one solve per CU (768 threads: three waves, up to 168 VGPRs; 512
// threads: two, up to 256), except 256 threads with two nodes per thread,
// which runs three solves per CU (three waves, up to 168 VGPRs); the other
// one-node shapes are left to the compiler.
// One RVE per wave (W > 0, W RVEs per workgroup of 64 W threads): two
// waves per SIMD (up to 256 VGPRs for three or four slots).
template <int B, int NPT, int RM, int W>
constexpr int reg_waves_min() {
  return W > 0 ? 2
         : RM == RM_TABLE ? 2 : NPT == 1 ? (B <= 512 ? 4 : 1) : (B == 768 || B == 256 ? 3 : 2);
}
template <int B, int NPT, int RM, int W>
constexpr int reg_waves_max() {
  return W > 0 || RM == RM_TABLE || NPT > 1 ? reg_waves_min<B, NPT, RM, W>() : 8;
}

// Orders the LDS accesses of one wave: its LDS operations execute in issue
// order, so only the compiler's code motion has to be fenced.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// W = 0: one RVE per workgroup of B threads.  W > 0 (B = 64 W): each wave
// solves its own small RVE in its own LDS region (a.wave_lds bytes, buffer
// stride kWaveBufStride), synchronising only within the wave, so a
// workgroup carries W RVEs and no workgroup barrier.
template <int B, int NPT, int RM, int W>
__global__ __launch_bounds__(B)
__attribute__((amdgpu_waves_per_eu(reg_waves_min<B, NPT, RM, W>(), reg_waves_max<B, NPT, RM, W>())))
void solve32r(SolveArgs a) {
  constexpr bool LIN = RM == RM_LIN;
  constexpr bool WM = W > 0;
  constexpr int TB = WM ? kWave : B;  // threads per RVE
  constexpr int BS = WM ? kWaveBufStride : kRegBufStride;
  constexpr int NW = TB / kWave;
  static_assert(!WM || B == W * kWave, "one RVE per wave");
  extern __shared__ __align__(16) unsigned char smem[];
  const int wave = WM ? (int)(threadIdx.x / kWave) : 0;
  const int tid = WM ? (int)(threadIdx.x & (kWave - 1)) : (int)threadIdx.x;
  auto sync = [&]() __attribute__((always_inline)) {
    if constexpr (WM) wave_sync();
    else __syncthreads();
  };
  // min of a and sum of b over the RVE's threads (every thread receives both)
  double *red = nullptr;
  auto red_min_sum = [&](double &x, double &y) __attribute__((always_inline)) {
    if constexpr (WM) {
      x = wave_min(x);
      y = wave_sum(y);
    } else {
      block_min_sum<B>(x, y, red);
    }
  };
  // step chunk of this workgroup (ramp mode): chunk `chunk` of the group's
  // solve `unit` (no chunks with one RVE per wave)
  int unit = WM ? (int)blockIdx.x * W + wave : (int)blockIdx.x, chunk = 0;
  if (WM && unit >= a.group_n) return;  // the last workgroup's spare waves
  const bool chunked = !WM && a.n_chunks > 1;
  if (chunked) {
    chunk = unit / a.group_n;
    unit -= chunk * a.group_n;
  }
  const bool last_chunk = !chunked || chunk == a.n_chunks - 1;
  const int sid = a.order ? a.order[unit] : unit;
  if (__builtin_amdgcn_groupstaticsize() != 0) {  // the LDS addressing below needs base 0
    if (tid == 0 && last_chunk) {
      for (int i = 0; i < 6; ++i) a.sigma[6 * (int64_t)sid + i] = __builtin_nan("");
      a.status[sid] = MF_STATUS_NAN;
    }
    return;
  }
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int64_t nb = R.node_off;

  // this RVE's LDS region (dynamic LDS starts at 0, so `region` is also
  // its byte address)
  const int region = WM ? wave * a.wave_lds : 0;
  unsigned char *const lds = smem + region;
  size_t off = 0;
  // displacement hi, two buffers at a fixed stride: the buffer of a step is a
  // compile-time constant, so it folds into the LDS loads' immediate offset
  float4 *pos = reinterpret_cast<float4 *>(lds);  // [2][BS]
  off += sizeof(float4) * (BS + Np);
  red = reinterpret_cast<double *>(lds + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(lds + off);  // [Nbp][6] u_b0, g
  off += sizeof(double) * 6 * Nbp;
  float4 *rtab = reinterpret_cast<float4 *>(lds + off);  // packed reaction table

  const int32_t *gw = a.ell_w + R.ell_off;
  const float *gil = a.ell_il + R.ell_off;
  const double *gl0 = a.ell_l0 + R.ell_off;
  const double *X = a.X + 3 * nb;
  if (tid < R.n_react) pack_react_table(a.reacts[R.react_off + tid], rtab + kReactTabStride * tid);
  const double EA_half = 0.5 * a.reacts[R.react_off].E * a.reacts[R.react_off].A;
  // RM_MIX2: the nonlinear reaction's constants (uniform)
  ReactDev nlq;
  if (RM == RM_MIX2) {
    const int k = (R.n_react > 1 && a.reacts[R.react_off + 1].type == MF_REACTION_NONLINEAR) ? 1 : 0;
    nlq = a.reacts[R.react_off + k];
  }

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;
  const double inv_S = 1.0 / a.n_steps;

  auto init_node = [&](int r, double (&u)[3]) {
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
    for (int k = 0; k < 3; ++k) u[k] = ramp ? up[k] : inc[k] + up[k];
    const float4 p = make_float4((float)u[0], (float)u[1], (float)u[2], 0.0f);
    pos[r] = p;
    pos[BS + r] = p;
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (!ramp) inc[k] = 0.0;
      bnd_set(bst + 6 * b, u, inc);
    }
  };
  for (int rb = tid; rb < N; rb += TB) {
    double u[3];
    init_node(rb, u);
  }

  const int S = R.n_split;
  const bool split = tid < 2 * S;  // slot 0 only
  int rr[NPT];
  bool own[NPT], writer[NPT];
  // imr: 1/m, or m in the linear mode, whose forces come scaled by 1/m
  float uh[NPT][3], ul[NPT][3], vr[NPT][3], imr[NPT];
  int offs[NPT][kLaneDeg];
  float cf[NPT][kLaneDeg], d0x[NPT][kLaneDeg], d0y[NPT][kLaneDeg], d0z[NPT][kLaneDeg];
  int Ds[NPT];
  unsigned nlm[NPT];
#pragma unroll
  for (int sl = 0; sl < NPT; ++sl) {
    const int r = sl == 0 ? (split ? tid >> 1 : tid - S) : sl * TB - S + tid;
    rr[sl] = r;
    own[sl] = r < Nf;
    writer[sl] = own[sl] && (sl > 0 || !split || !(tid & 1));
    imr[sl] = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) uh[sl][d] = ul[sl][d] = vr[sl][d] = 0.0f;
    int j0 = 0, dg = 0, base = 0, cnt = 0;
    if (own[sl]) {
      const int64_t g = nb + r;
      double xp[3], up[3];
      for (int d = 0; d < 3; ++d) {
        xp[d] = a.x_in[3 * g + d];
        up[d] = a.u_in[3 * g + d];
      }
      for (int k = 0; k < 3; ++k) {
        const double inc = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
                           (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
                           (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
        const double u = ramp ? up[k] : inc + up[k];
        uh[sl][k] = (float)u;
        ul[sl][k] = (float)(u - (double)uh[sl][k]);
      }
      imr[sl] = (float)(1.0 / a.mass[g]);
      if (LIN) imr[sl] = (float)a.mass[g];  // linear: forces come scaled by 1/m
      dg = a.deg[g];
      base = a.slice_base[R.slice_off + r / kWave] + (r & (kWave - 1));
      if (sl == 0 && split) {
        const int h = (dg + 1) >> 1;
        j0 = (tid & 1) ? h : 0;
        cnt = (tid & 1) ? dg - h : h;
      } else {
        cnt = dg;
      }
    }
    const int self = region + (own[sl] ? r : 0) * (int)sizeof(float4);
#pragma unroll
    for (int j = 0; j < kLaneDeg; ++j) {
      offs[sl][j] = self;
      cf[sl][j] = 0.0f;
      d0x[sl][j] = d0y[sl][j] = d0z[sl][j] = 0.0f;
      if (j < cnt) {
        const int idx = base + (j0 + j) * kWave;
        const int32_t wv = gw[idx];
        const int w = wv & 0xFFFFFF;
        offs[sl][j] = (region + w * (int)sizeof(float4)) | (RM == RM_TABLE ? (wv & 0xFF000000) : 0);
        const double l0 = gl0[idx];
        if (RM == RM_LIN) {
          // EA/(2 l0^3 m): the gathered force is the acceleration
          cf[sl][j] = (float)(EA_half / (l0 * l0 * l0 * a.mass[nb + r]));
        } else if (RM == RM_MIX2) {
          const ReactDev &q = a.reacts[R.react_off + ((uint32_t)wv >> 24)];
          cf[sl][j] = q.type == MF_REACTION_NONLINEAR ? -gil[idx]
                                                      : (float)(0.5 * q.E * q.A / (l0 * l0 * l0));
        } else {
          cf[sl][j] = gil[idx];
        }
        d0x[sl][j] = (float)(X[3 * w] - X[3 * r]);
        d0y[sl][j] = (float)(X[3 * w + 1] - X[3 * r + 1]);
        d0z[sl][j] = (float)(X[3 * w + 2] - X[3 * r + 2]);
      }
    }
    // wave-uniform trip count = largest per-lane count in the wave
    int D = cnt;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) D = max(D, __shfl_xor(D, o, kWave));
    Ds[sl] = __builtin_amdgcn_readfirstlane(D);
    // RM_MIX2: bit j set when a lane of the wave has a nonlinear fiber at
    // entry j (nodes list their nonlinear fibers first and a wave's nodes
    // have similar nonlinear counts, so the high entries are mostly linear
    // for the whole wave and skip the toe law)
    unsigned m = 0;
    if (RM == RM_MIX2) {
#pragma unroll
      for (int j = 0; j < kLaneDeg; ++j)
        m |= (__builtin_amdgcn_ballot_w64(cf[sl][j] < 0.0f) != 0 ? 1u : 0u) << j;
    }
    nlm[sl] = __builtin_amdgcn_readfirstlane(m);
  }
  // a later chunk continues from the state its predecessor saved
  float *cs = chunked ? a.cstate + (int64_t)unit * a.cstride : nullptr;
  float h_saved = 0.0f, res_saved = 0.0f;
  if (chunk > 0) {
    if (tid == 0) {
      bool ok = false;
      for (int spin = 0; spin < kChunkSpinLimit; ++spin) {
        if (__hip_atomic_load(a.cflag + sid, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= chunk) {
          ok = true;
          break;
        }
        __builtin_amdgcn_s_sleep(127);
      }
      if (!ok) cs[NPT * 9 * B + 1] = __builtin_nanf("");  // reported as NaN at the end
    }
    sync();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every lane sees the saved state
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uh[sl][d] = cs[(sl * 9 + d) * B + tid];
        ul[sl][d] = cs[(sl * 9 + 3 + d) * B + tid];
        vr[sl][d] = cs[(sl * 9 + 6 + d) * B + tid];
      }
    h_saved = cs[NPT * 9 * B];
    res_saved = cs[NPT * 9 * B + 1];
  }
  sync();

  const float inv_c = (float)(1.0 / R.c_sound);
  const float eta = (float)(a.eta_rve ? a.eta_rve[rv] : a.eta);
  int cur = 0;
  // internal force of a slot's node (both halves for a split node).  The
  // wave-uniform trip count selects a straight-line body, so the
  // neighbour-row loads issue ahead of the arithmetic.
  auto node_force = [&](int sl, auto BUFC, float (&f)[3], bool any_split)
                        __attribute__((always_inline)) {
    constexpr int buf = decltype(BUFC)::value;
    // LDS address of the buffer: dynamic LDS starts at 0 (the kernel has no
    // static LDS, checked on entry), so neighbour offset + buffer constant is
    // the whole address and the constant becomes the load's immediate offset
    const lds_uchar *P = reinterpret_cast<const lds_uchar *>(
        static_cast<uintptr_t>(buf * BS * (int)sizeof(float4)));
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    auto body = [&](auto KC) {
      constexpr int K = decltype(KC)::value;
      float4 q[K];
#pragma unroll
      for (int j = 0; j < K; ++j)
      {
        const v4f v = *reinterpret_cast<const lds_v4f *>(
            P + (RM == RM_TABLE ? (offs[sl][j] & 0xFFFFFF) : offs[sl][j]));
        q[j] = make_float4(v.x, v.y, v.z, 0.0f);
      }
#pragma unroll
      for (int j = 0; j < K; ++j) {
        // both ends as their FP32-rounded displacements (the low parts only
        // keep the accumulation exact)
        const float gx = q[j].x - uh[sl][0];
        const float gy = q[j].y - uh[sl][1];
        const float gz = q[j].z - uh[sl][2];
        const float dx = d0x[sl][j] + gx, dy = d0y[sl][j] + gy, dz = d0z[sl][j] + gz;
        // l^2 - l0^2 = (d0 + d) . delta
        const float sq = dot3f(d0x[sl][j] + dx, d0y[sl][j] + dy, d0z[sl][j] + dz, gx, gy, gz);
        float t;
        if (RM == RM_LIN) {
          t = cf[sl][j] * sq;
        } else if (RM == RM_MIX2) {
          // a wave with a nonlinear fiber at entry j evaluates both laws
          // on every lane and selects
          const float c = cf[sl][j];
          const float il = -c;
          if (nlm[sl] & (1u << j))
            t = c < 0.0f ? toe_ext_t(nlq, sq * (il * il), il) : c * sq;
          else
            t = c * sq;
        } else {
          const float il = cf[sl][j];
          t = react32_tab(rtab, (uint32_t)offs[sl][j] >> 24, sq * (il * il), il);
        }
        fx -= dx * t;
        fy -= dy * t;
        fz -= dz * t;
      }
    };
    switch (Ds[sl]) {
#define MF_DEG_CASE(k) \
  case k:              \
    body(std::integral_constant<int, k>{}); \
    break;
      MF_DEG_CASE(1) MF_DEG_CASE(2) MF_DEG_CASE(3)
      MF_DEG_CASE(4) MF_DEG_CASE(5) MF_DEG_CASE(6)
#undef MF_DEG_CASE
      default:
        break;
    }
    // combine the halves of split nodes: lanes 2i and 2i+1 swap partials
    if (any_split) {
      const float px = __builtin_amdgcn_update_dpp(0.0f, fx, 0xB1, 0xF, 0xF, false);
      const float py = __builtin_amdgcn_update_dpp(0.0f, fy, 0xB1, 0xF, 0xF, false);
      const float pz = __builtin_amdgcn_update_dpp(0.0f, fz, 0xB1, 0xF, 0xF, false);
      if (split) {
        fx += px;
        fy += py;
        fz += pz;
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  // min fiber length^2 over all fibers (every dt refresh)
  auto min_length2 = [&](int buf) {
    const float4 *P = pos + buf * BS;
    float lmin = 3.0e38f;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += TB) {
      const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
      const float4 p1 = P[n1], p2 = P[n2];
      const float x1 = (float)(X[3 * n2] - X[3 * n1]) + (p2.x - p1.x);
      const float x2 = (float)(X[3 * n2 + 1] - X[3 * n1 + 1]) + (p2.y - p1.y);
      const float x3 = (float)(X[3 * n2 + 2] - X[3 * n1 + 2]) + (p2.z - p1.z);
      lmin = fminf(lmin, dot3f(x1, x2, x3, x1, x2, x3));
    }
    return (double)lmin;
  };
  const bool wave_split = __builtin_amdgcn_readfirstlane(tid & ~(kWave - 1)) < 2 * S;

  // The closing half kick of step k and the opening half kick of step k+1
  // use the same acceleration: v_{k+1/2} = v_{k-1/2} + (h_k + h_{k+1}) a_k,
  // applied as soon as a_k and h_{k+1} are known (v_{-1/2} = 0, h_{-1} = 0),
  // so no acceleration is carried across the barrier.
  const float crit_half = (float)(a.crit * 0.5);
  float h_prev = h_saved;
  float fint[NPT][3];
  // acceleration of a slot's node: linear forces already carry 1/m
  auto accel = [&](int sl, int d) __attribute__((always_inline)) {
    return LIN ? fint[sl][d] : fint[sl][d] * imr[sl];
  };
  if (chunk == 0) {  // a later chunk's velocities already hold this kick
    double resid_local = 0.0;
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl)
      if (own[sl]) node_force(sl, std::integral_constant<int, 0>{}, fint[sl], true);
    double lmin2 = min_length2(cur);
    red_min_sum(lmin2, resid_local);
    const float h = crit_half * (sqrtf((float)lmin2) * inv_c);
    const float hk = h_prev + h;
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl)
      if (own[sl]) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[sl][d] = vr[sl][d] + hk * (-accel(sl, d) - eta * vr[sl][d]);
      }
    h_prev = h;
  }

  int step = chunked ? chunk * a.chunk_steps : 0;
  const int step_end = last_chunk ? a.n_steps : step + a.chunk_steps;
  double residual = chunk > 0 ? (double)res_saved : 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  // one step whose positions go to buffer NXT (a compile-time constant);
  // returns true when the solve (or its chunk) is done
  auto one_step = [&](auto NXTC) __attribute__((always_inline)) -> bool {
    constexpr int nxt = decltype(NXTC)::value;
    const float h2 = 2.0f * h_prev;
    float4 *Pn = pos + nxt * BS;
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl) {
      if (own[sl]) {
        float pc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          // u (hi, lo) += 2h v with Fast2Sum: exact when |u_hi| >= |increment|;
          // otherwise u is within a few increments of zero and the
          // uncompensated rounding is below eps * |increment|.
          const float y = ul[sl][d] + h2 * vr[sl][d];
          const float sh = uh[sl][d] + y;
          ul[sl][d] = y - (sh - uh[sl][d]);
          uh[sl][d] = sh;
          pc[d] = sh;
        }
        if (writer[sl]) Pn[rr[sl]] = make_float4(pc[0], pc[1], pc[2], 0.0f);
      }
    }
    if (ramp && tid < Nb) {  // waves without boundary nodes skip the amplitude
      const double alpha = ramp_alpha_step(step + 1, inv_S, a.amplitude);
      for (int b = tid; b < Nb; b += TB) {
        float ub[3];
        bnd_step(bst + 6 * b, alpha, ub);
        Pn[Nf + b] = make_float4(ub[0], ub[1], ub[2], 0.0f);
      }
    }
    sync();
    cur = nxt;
    const bool refresh = (step % freq) == 0;
    double rsum = 0.0;
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl) {
      if (own[sl]) {
        node_force(sl, NXTC, fint[sl], sl == 0 && wave_split);
        if (refresh && writer[sl]) {
          // the residual is |f|: linear forces are scaled back by m
          const float m = LIN ? imr[sl] : 1.0f;
          const float f0 = fint[sl][0] * m, f1 = fint[sl][1] * m, f2 = fint[sl][2] * m;
          rsum += (double)dot3f(f0, f1, f2, f0, f1, f2);
        }
      }
    }
    float h = h_prev;
    if (refresh) {
      double lm = min_length2(nxt);
      red_min_sum(lm, rsum);
      h = crit_half * (sqrtf((float)lm) * inv_c);
      residual = sqrt(rsum);
    }
    const float hk = h_prev + h;
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl)
      if (own[sl]) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[sl][d] = vr[sl][d] + hk * (-accel(sl, d) - eta * vr[sl][d]);
      }
    h_prev = h;
    ++step;
    if (ramp) return step >= step_end;
    if (!(residual > a.tol)) return true;
    if (step >= a.max_steps) {
      status = MF_STATUS_STEP_CAP;
      return true;
    }
    return false;
  };
  // the positions start in buffer 0; steps alternate buffers 1, 0, 1, ...
  // (every step writes all positions of its buffer before reading it, so a
  // later chunk may start on either buffer)
  for (;;) {
    if (one_step(std::integral_constant<int, 1>{})) break;
    if (one_step(std::integral_constant<int, 0>{})) break;
  }
  if (!last_chunk) {
    // save the lane state, publish it, raise the solve's chunk count
#pragma unroll
    for (int sl = 0; sl < NPT; ++sl)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        cs[(sl * 9 + d) * B + tid] = uh[sl][d];
        cs[(sl * 9 + 3 + d) * B + tid] = ul[sl][d];
        cs[(sl * 9 + 6 + d) * B + tid] = vr[sl][d];
      }
    if (tid == 0) {
      cs[NPT * 9 * B] = h_prev;
      cs[NPT * 9 * B + 1] = (float)residual;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    sync();
    if (tid == 0)
      __hip_atomic_store(a.cflag + sid, chunk + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (chunked && tid == 0) a.cflag[sid] = 0;  // the next launch starts from zero
  if (residual != residual) status = MF_STATUS_NAN;

  // finalize: low parts into the spare buffer, boundary gathers in FP32 with
  // exactly rounded reference vectors, stress in FP64
  float4 *Ph = pos + cur * BS;
  float4 *Pl = pos + (cur ^ 1) * BS;
  sync();
#pragma unroll
  for (int sl = 0; sl < NPT; ++sl)
    if (writer[sl]) Pl[rr[sl]] = make_float4(ul[sl][0], ul[sl][1], ul[sl][2], 0.0f);
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += TB) {
    double ub[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    const float4 hb = make_float4((float)ub[0], (float)ub[1], (float)ub[2], 0.0f);
    Ph[Nf + b] = hb;
    Pl[Nf + b] = make_float4((float)(ub[0] - (double)hb.x), (float)(ub[1] - (double)hb.y),
                             (float)(ub[2] - (double)hb.z), 0.0f);
  }
  sync();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  for (int b = tid; b < Nb; b += TB) {
    const int rb = Nf + b;
    const int64_t gr = nb + rb;
    double ub[3], xb[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    for (int d = 0; d < 3; ++d) xb[d] = X[3 * rb + d] + ub[d];
    const float4 hb = Ph[rb], lb = Pl[rb];
    const int k = rb / kWave;
    const int base = a.slice_base[R.slice_off + k] + (rb & (kWave - 1));
    const int dg = a.deg[gr];
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const int w = wv & 0xFFFFFF;
      const float4 qh = Ph[w], ql = Pl[w];
      const float ex = (float)(X[3 * w] - X[3 * rb]);
      const float ey = (float)(X[3 * w + 1] - X[3 * rb + 1]);
      const float ez = (float)(X[3 * w + 2] - X[3 * rb + 2]);
      const float gx = (qh.x - hb.x) + (ql.x - lb.x);
      const float gy = (qh.y - hb.y) + (ql.y - lb.y);
      const float gz = (qh.z - hb.z) + (ql.z - lb.z);
      const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
      const float sq = dot3f(ex + dx, ey + dy, ez + dz, gx, gy, gz);
      float t;
      if (LIN) {
        const double l0 = gl0[idx];
        t = (float)(EA_half / (l0 * l0 * l0)) * sq;
      } else {
        const float il = gil[idx];
        t = react32_tab(rtab, (uint32_t)wv >> 24, sq * (il * il), il);
      }
      f1 -= dx * t;
      f2 -= dy * t;
      f3 -= dz * t;
    }
    const double g1 = f1, g2 = f2, g3 = f3;
    st[0] += xb[0] * g1;
    st[1] += xb[1] * g2;
    st[2] += xb[2] * g3;
    st[3] += 0.5 * fma(xb[1], g3, xb[2] * g2);
    st[4] += 0.5 * fma(xb[0], g3, xb[2] * g1);
    st[5] += 0.5 * fma(xb[0], g2, xb[1] * g1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * gr + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * gr + d] = ub[d];
  }
#pragma unroll
  for (int sl = 0; sl < NPT; ++sl)
    if (writer[sl]) {
      const int64_t g = nb + rr[sl];
      for (int d = 0; d < 3; ++d) {
        const double u = (double)uh[sl][d] + (double)ul[sl][d];
        if (flags & kWriteX) a.x_out[3 * g + d] = X[3 * rr[sl] + d] + u;
        if (flags & kWriteU) a.u_out[3 * g + d] = u;
      }
    }
  if constexpr (WM) {
#pragma unroll
    for (int i = 0; i < 6; ++i) st[i] = wave_sum(st[i]);
  } else {
    block_sum<B, 6>(st, red);
  }
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
// FP32c fiber-parallel kernel
// ============================================================================
// Lanes own fibers: each fiber is evaluated once per step and publishes its
// two end forces, -t d to end a and +t d to end b, into LDS slots laid out as
// the sliced ELL (the entry of the fiber in each end's incidence list; ends
// at boundary nodes write a spare slot of their lane, so they never contend
// for one address).  Free nodes then sum their own
// slots, which sit lane-consecutive, kick and drift.  Two barriers per step:
// positions published, end forces published.  The per-fiber arithmetic is
// the register kernel's (both ends as FP32-rounded displacements, exactly
// rounded reference vectors), done once per fiber instead of once per end.
// Slots that no fiber owns (ELL padding) stay zero, so nodes sum all entries
// of their slice width unmasked.
template <int B, int FPL, int NPL, int RM>
__global__ __launch_bounds__(B) void solve32f(SolveArgs a) {
  constexpr bool LIN = RM == RM_LIN;
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int ne = R.ell_free_entries;
  const int nslot = max(ne + kWave, Np);
  const int64_t nb = R.node_off;

  size_t off = 0;
  float4 *pos = reinterpret_cast<float4 *>(smem);  // [Np] displacement hi
  off += sizeof(float4) * Np;
  float4 *fsl = reinterpret_cast<float4 *>(smem + off);  // [nslot] end forces
  off += sizeof(float4) * nslot;
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);  // [Nbp][6] u_b0, g
  off += sizeof(double) * 6 * Nbp;
  float4 *rtab = reinterpret_cast<float4 *>(smem + off);  // packed reaction table

  const int32_t *gw = a.ell_w + R.ell_off;
  const float *gil = a.ell_il + R.ell_off;
  const double *gl0 = a.ell_l0 + R.ell_off;
  const double *X = a.X + 3 * nb;
  if (tid < R.n_react) pack_react_table(a.reacts[R.react_off + tid], rtab + kReactTabStride * tid);
  const double EA_half = 0.5 * a.reacts[R.react_off].E * a.reacts[R.react_off].A;
  ReactDev nlq;
  if (RM == RM_MIX2) {
    const int k = (R.n_react > 1 && a.reacts[R.react_off + 1].type == MF_REACTION_NONLINEAR) ? 1 : 0;
    nlq = a.reacts[R.react_off + k];
  }

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;
  const double inv_S = 1.0 / a.n_steps;

  // initial displacement: relax adds the affine increment (F - I) x to u,
  // ramp keeps u and sets the boundary increment aside
  auto init_u = [&](int r, double (&u)[3], double (&inc)[3]) {
    const int64_t g = nb + r;
    double xp[3], up[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = a.x_in[3 * g + d];
      up[d] = a.u_in[3 * g + d];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      inc[k] = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
               (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
               (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
      u[k] = ramp ? up[k] : inc[k] + up[k];
    }
  };
  for (int r = tid; r < N; r += B) {
    double u[3], inc[3];
    init_u(r, u, inc);
    pos[r] = make_float4((float)u[0], (float)u[1], (float)u[2], 0.0f);
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (!ramp) inc[k] = 0.0;
      bnd_set(bst + 6 * b, u, inc);
    }
  }
  for (int i = tid; i < nslot; i += B) fsl[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

  // free nodes: r = s * B + tid
  float uh[NPL][3], ul[NPL][3], vr[NPL][3], imr[NPL];
  int degn[NPL], Dn[NPL], sbn[NPL];
#pragma unroll
  for (int s = 0; s < NPL; ++s) {
    const int r = s * B + tid;
    imr[s] = 0.0f;
    degn[s] = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) uh[s][d] = ul[s][d] = vr[s][d] = 0.0f;
    if (r < Nf) {
      double u[3], inc[3];
      init_u(r, u, inc);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uh[s][d] = (float)u[d];
        ul[s][d] = (float)(u[d] - (double)uh[s][d]);
      }
      imr[s] = (float)(1.0 / a.mass[nb + r]);
      degn[s] = a.deg[nb + r];
    }
    // lane 0 of the wave names the slice; waves past Nf never gather
    const int k = r / kWave;
    Dn[s] = __builtin_amdgcn_readfirstlane(r < Nf ? a.slice_D[R.slice_off + k] : 0);
    sbn[s] = __builtin_amdgcn_readfirstlane(r < Nf ? a.slice_base[R.slice_off + k] : 0);
  }

  // fibers: e = s * B + tid; byte offsets of the end positions (reaction id
  // in the top byte in table mode) and of the end-force slots
  int pa[FPL], pb[FPL], fa[FPL], fb[FPL];
  float d0x[FPL], d0y[FPL], d0z[FPL], cf[FPL];
  const int32_t *fs = a.fstream + 4 * R.fib_off;
#pragma unroll
  for (int s = 0; s < FPL; ++s) {
    const int e = s * B + tid;
    pa[s] = pb[s] = 0;
    fa[s] = fb[s] = (ne + lane) * (int)sizeof(float4);
    d0x[s] = d0y[s] = d0z[s] = 0.0f;
    cf[s] = 0.0f;
    if (e < M) {
      const int4 q = *reinterpret_cast<const int4 *>(fs + 4 * e);
      pa[s] = q.x * (int)sizeof(float4);
      pb[s] = q.y * (int)sizeof(float4);
      if (q.z >= 0) fa[s] = q.z * (int)sizeof(float4);
      if (q.w >= 0) fb[s] = q.w * (int)sizeof(float4);
      d0x[s] = (float)(X[3 * q.y] - X[3 * q.x]);
      d0y[s] = (float)(X[3 * q.y + 1] - X[3 * q.x + 1]);
      d0z[s] = (float)(X[3 * q.y + 2] - X[3 * q.x + 2]);
      if (q.z >= 0) {  // fibers with two boundary ends only set the time step
        const double l0 = gl0[q.z];
        const int32_t wv = gw[q.z];
        if (RM == RM_LIN) {
          cf[s] = (float)(EA_half / (l0 * l0 * l0));
        } else if (RM == RM_MIX2) {
          const ReactDev &rq = a.reacts[R.react_off + ((uint32_t)wv >> 24)];
          cf[s] = rq.type == MF_REACTION_NONLINEAR ? -gil[q.z]
                                                   : (float)(0.5 * rq.E * rq.A / (l0 * l0 * l0));
        } else {
          cf[s] = gil[q.z];
          pa[s] |= wv & 0xFF000000;
        }
      }
    }
  }
  __syncthreads();

  const float inv_c = (float)(1.0 / R.c_sound);
  const float eta = (float)(a.eta_rve ? a.eta_rve[rv] : a.eta);

  // end forces of every owned fiber into the slots; squared current length
  // minimum over the owned fibers when asked
  // waves whose fibers of a slot all lie past M skip that slot
  // nlw: a fiber of the wave's slot is nonlinear (the stream lists an RVE's
  // nonlinear fibers first, so most waves hold one law)
  bool live[FPL], nlw[FPL];
#pragma unroll
  for (int s = 0; s < FPL; ++s) {
    live[s] = __builtin_amdgcn_readfirstlane(s * B + (tid & ~(kWave - 1))) < M;
    nlw[s] = __builtin_amdgcn_readfirstlane(
                 __builtin_amdgcn_ballot_w64(cf[s] < 0.0f) != 0 ? 1 : 0) != 0;
  }
  auto fiber_phase = [&](bool want_min) {
    const unsigned char *P = reinterpret_cast<const unsigned char *>(pos);
    unsigned char *Fs = reinterpret_cast<unsigned char *>(fsl);
    float lmin = 3.0e38f;
    float4 qa[FPL], qb[FPL];
#pragma unroll
    for (int s = 0; s < FPL; ++s)
      if (live[s]) {
        qa[s] = ld16(reinterpret_cast<const float4 *>(
            P + (RM == RM_TABLE ? (pa[s] & 0xFFFFFF) : pa[s])));
        qb[s] = ld16(reinterpret_cast<const float4 *>(P + pb[s]));
      }
#pragma unroll
    for (int s = 0; s < FPL; ++s) {
      if (!live[s]) continue;
      const float gx = qb[s].x - qa[s].x;
      const float gy = qb[s].y - qa[s].y;
      const float gz = qb[s].z - qa[s].z;
      const float dx = d0x[s] + gx, dy = d0y[s] + gy, dz = d0z[s] + gz;
      // l^2 - l0^2 = (d0 + d) . delta
      const float sq = dot3f(d0x[s] + dx, d0y[s] + dy, d0z[s] + dz, gx, gy, gz);
      float t;
      if (RM == RM_LIN) {
        t = cf[s] * sq;
      } else if (RM == RM_MIX2) {
        const float c = cf[s];
        const float il = -c;
        if (nlw[s])
          t = c < 0.0f ? toe_ext_t(nlq, sq * (il * il), il) : c * sq;
        else
          t = c * sq;
      } else {
        const float il = cf[s];
        t = react32_tab(rtab, (uint32_t)pa[s] >> 24, sq * (il * il), il);
      }
      if (want_min && s * B + tid < M) lmin = fminf(lmin, dot3f(dx, dy, dz, dx, dy, dz));
      const float fx = dx * t, fy = dy * t, fz = dz * t;
      *reinterpret_cast<float4 *>(Fs + fa[s]) = make_float4(-fx, -fy, -fz, 0.0f);
      *reinterpret_cast<float4 *>(Fs + fb[s]) = make_float4(fx, fy, fz, 0.0f);
    }
    return lmin;
  };
  // internal force of a free node: the sum of its slots.  The wave-uniform
  // slice width selects a straight-line body so the loads issue together.
  auto node_force = [&](int s, float (&f)[3]) {
    const float4 *Q = fsl + sbn[s] + lane;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    // loads in batches of at most 6 rows (bounded registers)
    auto batch = [&](auto J0C, auto KC) {
      constexpr int J0 = decltype(J0C)::value, K = decltype(KC)::value;
      float4 q[K];
#pragma unroll
      for (int j = 0; j < K; ++j) q[j] = ld16(Q + (J0 + j) * kWave);
#pragma unroll
      for (int j = 0; j < K; ++j) {
        fx += q[j].x;
        fy += q[j].y;
        fz += q[j].z;
      }
    };
    auto body = [&](auto KC) {
      constexpr int K = decltype(KC)::value;
      batch(std::integral_constant<int, 0>{}, std::integral_constant<int, (K < 6 ? K : 6)>{});
      if constexpr (K > 6)
        batch(std::integral_constant<int, 6>{}, std::integral_constant<int, K - 6>{});
    };
    switch (Dn[s]) {
#define MF_DEG_CASE(k) \
  case k:              \
    body(std::integral_constant<int, k>{}); \
    break;
      MF_DEG_CASE(1) MF_DEG_CASE(2) MF_DEG_CASE(3) MF_DEG_CASE(4)
      MF_DEG_CASE(5) MF_DEG_CASE(6) MF_DEG_CASE(7) MF_DEG_CASE(8)
      MF_DEG_CASE(9) MF_DEG_CASE(10) MF_DEG_CASE(11) MF_DEG_CASE(12)
#undef MF_DEG_CASE
      default:
        for (int j = 0; j < Dn[s]; ++j) {
          const float4 q = ld16(Q + j * kWave);
          fx += q.x;
          fy += q.y;
          fz += q.z;
        }
        break;
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  auto wave_min_to = [&](float v, double *slot) {
    double m = wave_min((double)v);
    if (lane == 0) slot[(tid / kWave) * kRedStride] = m;
  };
  auto red_min = [&](const double *slot) {
    double m = slot[0];
#pragma unroll
    for (int k = 1; k < NW; ++k) m = fmin(m, slot[k * kRedStride]);
    return m;
  };

  // The closing half kick of step k and the opening half kick of step k+1
  // use the same acceleration: v_{k+1/2} = v_{k-1/2} + (h_k + h_{k+1}) a_k,
  // applied as soon as a_k and h_{k+1} are known (v_{-1/2} = 0, h_{-1} = 0).
  const float crit_half = (float)(a.crit * 0.5);
  float h_prev = 0.0f;
  {
    const float lm = fiber_phase(true);
    wave_min_to(lm, red);
    __syncthreads();
    const float h = crit_half * (sqrtf((float)red_min(red)) * inv_c);
#pragma unroll
    for (int s = 0; s < NPL; ++s)
      if (s * B + tid < Nf) {
        float f[3];
        node_force(s, f);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[s][d] = vr[s][d] + h * (-f[d] * imr[s] - eta * vr[s][d]);
      }
    h_prev = h;
  }

  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  for (;;) {
    const float h2 = 2.0f * h_prev;
    // drift: u (hi, lo) += 2h v with Fast2Sum, exact when |u_hi| >= |increment|;
    // otherwise u is within a few increments of zero and the uncompensated
    // rounding is below eps * |increment|.
#pragma unroll
    for (int s = 0; s < NPL; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
        float pc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float y = ul[s][d] + h2 * vr[s][d];
          const float sh = uh[s][d] + y;
          ul[s][d] = y - (sh - uh[s][d]);
          uh[s][d] = sh;
          pc[d] = sh;
        }
        pos[r] = make_float4(pc[0], pc[1], pc[2], 0.0f);
      }
    }
    if (ramp && tid < Nb) {  // waves without boundary nodes skip the amplitude
      const double alpha = ramp_alpha_step(step + 1, inv_S, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
        float ub[3];
        bnd_step(bst + 6 * b, alpha, ub);
        pos[Nf + b] = make_float4(ub[0], ub[1], ub[2], 0.0f);
      }
    }
    __syncthreads();  // positions published; every slot read of the last step done
    const bool refresh = (step % freq) == 0;
    const float lm = fiber_phase(refresh);
    if (refresh) wave_min_to(lm, red);
    __syncthreads();  // end forces published
    float h = h_prev;
    if (refresh) h = crit_half * (sqrtf((float)red_min(red)) * inv_c);
    const float hk = h_prev + h;
    double rsum = 0.0;
#pragma unroll
    for (int s = 0; s < NPL; ++s)
      if (s * B + tid < Nf) {
        float f[3];
        node_force(s, f);
        if (refresh) rsum += (double)dot3f(f[0], f[1], f[2], f[0], f[1], f[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[s][d] = vr[s][d] + hk * (-f[d] * imr[s] - eta * vr[s][d]);
      }
    if (refresh) {
      // residual in slots 1 of the scratch (slots 0 are read above)
      const double ws = wave_sum(rsum);
      if (lane == 0) red[(tid / kWave) * kRedStride + 1] = ws;
      __syncthreads();
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) sum += red[k * kRedStride + 1];
      residual = sqrt(sum);
    }
    h_prev = h;
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

  // finalize: low parts into the slot buffer, boundary gathers in FP32 with
  // exactly rounded reference vectors, stress in FP64
  float4 *Ph = pos;
  float4 *Pl = fsl;
  __syncthreads();  // every slot and scratch read is done
#pragma unroll
  for (int s = 0; s < NPL; ++s) {
    const int r = s * B + tid;
    if (r < Nf) Pl[r] = make_float4(ul[s][0], ul[s][1], ul[s][2], 0.0f);
  }
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += B) {
    double ub[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    const float4 hb = make_float4((float)ub[0], (float)ub[1], (float)ub[2], 0.0f);
    Ph[Nf + b] = hb;
    Pl[Nf + b] = make_float4((float)(ub[0] - (double)hb.x), (float)(ub[1] - (double)hb.y),
                             (float)(ub[2] - (double)hb.z), 0.0f);
  }
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  for (int b = tid; b < Nb; b += B) {
    const int rb = Nf + b;
    const int64_t gr = nb + rb;
    double ub[3], xb[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    for (int d = 0; d < 3; ++d) xb[d] = X[3 * rb + d] + ub[d];
    const float4 hb = Ph[rb], lb = Pl[rb];
    const int k = rb / kWave;
    const int base = a.slice_base[R.slice_off + k] + (rb & (kWave - 1));
    const int dg = a.deg[gr];
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const int w = wv & 0xFFFFFF;
      const float4 qh = Ph[w], ql = Pl[w];
      const float ex = (float)(X[3 * w] - X[3 * rb]);
      const float ey = (float)(X[3 * w + 1] - X[3 * rb + 1]);
      const float ez = (float)(X[3 * w + 2] - X[3 * rb + 2]);
      const float gx = (qh.x - hb.x) + (ql.x - lb.x);
      const float gy = (qh.y - hb.y) + (ql.y - lb.y);
      const float gz = (qh.z - hb.z) + (ql.z - lb.z);
      const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
      const float sq = dot3f(ex + dx, ey + dy, ez + dz, gx, gy, gz);
      float t;
      if (LIN) {
        const double l0 = gl0[idx];
        t = (float)(EA_half / (l0 * l0 * l0)) * sq;
      } else {
        const float il = gil[idx];
        t = react32_tab(rtab, (uint32_t)wv >> 24, sq * (il * il), il);
      }
      f1 -= dx * t;
      f2 -= dy * t;
      f3 -= dz * t;
    }
    const double g1 = f1, g2 = f2, g3 = f3;
    st[0] += xb[0] * g1;
    st[1] += xb[1] * g2;
    st[2] += xb[2] * g3;
    st[3] += 0.5 * fma(xb[1], g3, xb[2] * g2);
    st[4] += 0.5 * fma(xb[0], g3, xb[2] * g1);
    st[5] += 0.5 * fma(xb[0], g2, xb[1] * g1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * gr + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * gr + d] = ub[d];
  }
#pragma unroll
  for (int s = 0; s < NPL; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      const int64_t g = nb + r;
      for (int d = 0; d < 3; ++d) {
        const double u = (double)uh[s][d] + (double)ul[s][d];
        if (flags & kWriteX) a.x_out[3 * g + d] = X[3 * r + d] + u;
        if (flags & kWriteU) a.u_out[3 * g + d] = u;
      }
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
// FP32c fiber-streamed kernel (large RVEs)
// ============================================================================
// Each fiber is evaluated once per step by the lane that owns it (endpoint
// byte offsets packed in one register, reference vector and law coefficient
// in registers; streaming them from global memory each step measured
// slower) and its force t d is written to a 12-byte record in LDS, in
// fiber-stream order (lane-consecutive stores).  Each free node then sums
// +-force over its incidence entries, whose record offsets and signs stream
// from the ELL in global memory with a rolling prefetch.  Positions and
// forces are 12-byte records, so LDS holds about 37 bytes per node plus the
// boundary state: RVEs up to ~3700 nodes.  Arithmetic per fiber is the
// register kernel's (both ends as FP32-rounded displacements, exactly rounded
// reference vectors).
struct F3 {
  float x, y, z;
};

template <int B, int FPL, int NPL, int RM>
__global__ __launch_bounds__(B) void solve32s(SolveArgs a) {
  constexpr bool LIN = RM == RM_LIN;
  constexpr bool RR = B <= 768;
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const int rv = a.solve_rve[sid];
  const RveDev R = a.rves[rv];
  const int N = R.N, Nf = R.Nf, Nb = R.Nb, M = R.M;
  const int Np = (N + 1) & ~1, Nbp = (Nb + 1) & ~1;
  const int nff = ((max(M + 1, N)) + 1) & ~1;  // force records (+ zero record)
  const int64_t nb = R.node_off;

  size_t off = 0;
  F3 *pos = reinterpret_cast<F3 *>(smem);  // [Np] displacement hi
  off += sizeof(F3) * Np;
  F3 *ffr = reinterpret_cast<F3 *>(smem + off);  // [nff] fiber forces t d
  off += sizeof(F3) * nff;
  off = (off + 15) & ~size_t(15);  // 16-byte boundary records below
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);  // [Nbp][6] u_b0, g
  off += sizeof(double) * 6 * Nbp;
  off = (off + 15) & ~size_t(15);
  float4 *rtab = reinterpret_cast<float4 *>(smem + off);  // packed table (finalize)

  const int32_t *gw = a.ell_w + R.ell_off;
  const float *gil = a.ell_il + R.ell_off;
  const double *gl0 = a.ell_l0 + R.ell_off;
  const int32_t *gfr = a.ell_fref + R.ell_off;
  // the free nodes' force-record references as a raw buffer (loads past the
  // end return 0; every load here stays within the slices)
  const __amdgpu_buffer_rsrc_t rsfr = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<int32_t *>(gfr), 0, R.ell_free_entries * (int)sizeof(int32_t), kBufRsrcWord3);
  const double *X = a.X + 3 * nb;
  if (tid < R.n_react) pack_react_table(a.reacts[R.react_off + tid], rtab + kReactTabStride * tid);
  const double EA_half = 0.5 * a.reacts[R.react_off].E * a.reacts[R.react_off].A;
  ReactDev nlq;
  if (RM == RM_MIX2) {
    const int k = (R.n_react > 1 && a.reacts[R.react_off + 1].type == MF_REACTION_NONLINEAR) ? 1 : 0;
    nlq = a.reacts[R.react_off + k];
  }

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;
  const double inv_S = 1.0 / a.n_steps;

  auto init_u = [&](int r, double (&u)[3], double (&inc)[3]) {
    const int64_t g = nb + r;
    double xp[3], up[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = a.x_in[3 * g + d];
      up[d] = a.u_in[3 * g + d];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      inc[k] = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
               (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
               (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
      u[k] = ramp ? up[k] : inc[k] + up[k];
    }
  };
  for (int r = tid; r < N; r += B) {
    double u[3], inc[3];
    init_u(r, u, inc);
    pos[r] = F3{(float)u[0], (float)u[1], (float)u[2]};
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (!ramp) inc[k] = 0.0;
      bnd_set(bst + 6 * b, u, inc);
    }
  }
  for (int i = tid; i < nff; i += B) ffr[i] = F3{0.0f, 0.0f, 0.0f};

  // free nodes: r = s * B + tid
  float uh[NPL][3], ul[NPL][3], vr[NPL][3], imr[NPL];
  int Dn[NPL], sbn[NPL];
  // RR (up to 768 threads, whose register budget holds them): the first
  // kLaneDeg force-record references of each node slot in registers, two
  // 16-bit fields per register (record index, bit 15 set where the node is
  // the fiber's end a and takes -force)
  constexpr int kRw = (kLaneDeg + 1) / 2;
  uint32_t rpk[NPL][RR ? kRw : 1];
#pragma unroll
  for (int s = 0; s < NPL; ++s) {
    const int r = s * B + tid;
    imr[s] = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) uh[s][d] = ul[s][d] = vr[s][d] = 0.0f;
    if (r < Nf) {
      double u[3], inc[3];
      init_u(r, u, inc);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uh[s][d] = (float)u[d];
        ul[s][d] = (float)(u[d] - (double)uh[s][d]);
      }
      imr[s] = (float)(1.0 / a.mass[nb + r]);
    }
    const int k = r / kWave;
    Dn[s] = __builtin_amdgcn_readfirstlane(r < Nf ? a.slice_D[R.slice_off + k] : 0);
    sbn[s] = __builtin_amdgcn_readfirstlane(r < Nf ? a.slice_base[R.slice_off + k] : 0);
    if constexpr (RR) {
#pragma unroll
      for (int w = 0; w < kRw; ++w) rpk[s][w] = 0;
#pragma unroll
      for (int j = 0; j < kLaneDeg; ++j) {
        const int32_t fr = (r < Nf && j < Dn[s]) ? gfr[sbn[s] + lane + j * kWave] : M;
        const uint32_t field = ((uint32_t)fr & 0x7FFFu) | (fr < 0 ? 0x8000u : 0u);
        rpk[s][j / 2] |= field << (16 * (j & 1));
      }
    }
  }

  // fibers: e = s * B + tid in stream order; endpoint byte offsets a | b << 16
  int pab[FPL];
  float d0x[FPL], d0y[FPL], d0z[FPL], cf[FPL];
  // nlw: a fiber of the wave's slot is nonlinear (the stream lists an RVE's
  // nonlinear fibers first, so most waves hold one law)
  bool live[FPL], nlw[FPL];
  const int32_t *fs = a.fstream + 4 * R.fib_off;
  const float4 *fp = reinterpret_cast<const float4 *>(a.fparam) + R.fib_off;
#pragma unroll
  for (int s = 0; s < FPL; ++s) {
    const int e = s * B + tid;
    live[s] = __builtin_amdgcn_readfirstlane(s * B + (tid & ~(kWave - 1))) < M;
    pab[s] = 0;
    d0x[s] = d0y[s] = d0z[s] = cf[s] = 0.0f;
    if (e < M) {
      const int4 q = *reinterpret_cast<const int4 *>(fs + 4 * e);
      pab[s] = (q.x * (int)sizeof(F3)) | ((q.y * (int)sizeof(F3)) << 16);
      const float4 c = fp[e];
      d0x[s] = c.x;
      d0y[s] = c.y;
      d0z[s] = c.z;
      cf[s] = c.w;
    }
    nlw[s] = __builtin_amdgcn_readfirstlane(
                 __builtin_amdgcn_ballot_w64(cf[s] < 0.0f) != 0 ? 1 : 0) != 0;
  }
  __syncthreads();

  const float inv_c = (float)(1.0 / R.c_sound);
  const float eta = (float)(a.eta_rve ? a.eta_rve[rv] : a.eta);

  // forces t d of the owned fibers into their records; squared current
  // length minimum over the owned fibers when asked
  auto fiber_phase = [&](bool want_min) {
    const unsigned char *P = reinterpret_cast<const unsigned char *>(pos);
    float lmin = 3.0e38f;
#pragma unroll
    for (int s = 0; s < FPL; ++s) {
      if (!live[s]) continue;
      const float px = d0x[s], py = d0y[s], pz = d0z[s], pc = cf[s];
      const F3 qa = *reinterpret_cast<const F3 *>(P + (pab[s] & 0xFFFF));
      const F3 qb = *reinterpret_cast<const F3 *>(P + ((uint32_t)pab[s] >> 16));
      const float gx = qb.x - qa.x, gy = qb.y - qa.y, gz = qb.z - qa.z;
      const float dx = px + gx, dy = py + gy, dz = pz + gz;
      // l^2 - l0^2 = (d0 + d) . delta
      const float sq = dot3f(px + dx, py + dy, pz + dz, gx, gy, gz);
      float t;
      if (RM == RM_LIN) {
        t = pc * sq;
      } else {
        const float c = pc;
        const float il = -c;
        if (nlw[s])
          t = c < 0.0f ? toe_ext_t(nlq, sq * (il * il), il) : c * sq;
        else
          t = c * sq;
      }
      const int e = s * B + tid;
      if (e < M) {
        if (want_min) lmin = fminf(lmin, dot3f(dx, dy, dz, dx, dy, dz));
        ffr[e] = F3{dx * t, dy * t, dz * t};
      }
    }
    return lmin;
  };
  // internal force of a free node: -force of fibers where it is end a,
  // +force where it is end b, summed in list order.  RR: the first kLaneDeg
  // references come from registers and the wave-uniform slice width selects
  // a straight-line body, so the record loads issue together; entries past
  // kLaneDeg (high-degree slices only) load their references from global
  // memory four at a time.  Otherwise the references stream from global
  // memory one entry ahead.
  auto node_force = [&](int s, float (&f)[3]) {
    const int D = Dn[s];
    const int32_t *E = gfr + sbn[s] + lane;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if constexpr (RR) {
      // opaque to the compiler: decoding the references ahead of the step
      // loop would hold every decoded address and sign in registers
#pragma unroll
      for (int w = 0; w < kRw; ++w) asm volatile("" : "+v"(rpk[s][w]));
      auto body = [&](auto KC) {
        constexpr int K = decltype(KC)::value;
        F3 v[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const uint32_t field = (rpk[s][j / 2] >> (16 * (j & 1))) & 0xFFFFu;
          v[j] = ffr[field & 0x7FFFu];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const uint32_t field = (rpk[s][j / 2] >> (16 * (j & 1))) & 0xFFFFu;
          const float sg = __uint_as_float(0x3F800000u | ((field & 0x8000u) << 16));
          fx = fmaf(sg, v[j].x, fx);
          fy = fmaf(sg, v[j].y, fy);
          fz = fmaf(sg, v[j].z, fz);
        }
      };
      switch (D < kLaneDeg ? D : kLaneDeg) {
#define MF_DEG_CASE(k) \
  case k:              \
    body(std::integral_constant<int, k>{}); \
    break;
        MF_DEG_CASE(1) MF_DEG_CASE(2) MF_DEG_CASE(3)
        MF_DEG_CASE(4) MF_DEG_CASE(5) MF_DEG_CASE(6)
#undef MF_DEG_CASE
        default:
          break;
      }
      // four references per group load together (indices clamped to the
      // slice's last entry) and are then summed in list order
      const int vo = (sbn[s] + lane) * (int)sizeof(int32_t);
      for (int j = kLaneDeg; j < D; j += 4) {
        int32_t fr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          fr[q] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(
              rsfr, vo, min(j + q, D - 1) * (kWave * (int)sizeof(int32_t)), 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (j + q < D) {
            const F3 v = ffr[fr[q] & 0x7FFFFFFF];
            const float sg = fr[q] < 0 ? -1.0f : 1.0f;
            fx = fmaf(sg, v.x, fx);
            fy = fmaf(sg, v.y, fy);
            fz = fmaf(sg, v.z, fz);
          }
        }
      }
    } else if (D > 0) {
      int32_t nx = E[0];
      for (int j = 0; j < D; ++j) {
        const int32_t fr = nx;
        if (j + 1 < D) nx = E[(j + 1) * kWave];
        const F3 v = ffr[fr & 0x7FFFFFFF];
        const float sg = __int_as_float((fr & (int32_t)0x80000000) | 0x3F800000);
        fx = fmaf(sg, v.x, fx);
        fy = fmaf(sg, v.y, fy);
        fz = fmaf(sg, v.z, fz);
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  auto wave_min_to = [&](float v, double *slot) {
    double m = wave_min((double)v);
    if (lane == 0) slot[(tid / kWave) * kRedStride] = m;
  };
  auto red_min = [&](const double *slot) {
    double m = slot[0];
#pragma unroll
    for (int k = 1; k < NW; ++k) m = fmin(m, slot[k * kRedStride]);
    return m;
  };

  // The closing half kick of step k and the opening half kick of step k+1
  // use the same acceleration: v_{k+1/2} = v_{k-1/2} + (h_k + h_{k+1}) a_k,
  // applied as soon as a_k and h_{k+1} are known (v_{-1/2} = 0, h_{-1} = 0).
  const float crit_half = (float)(a.crit * 0.5);
  float h_prev = 0.0f;
  {
    const float lm = fiber_phase(true);
    wave_min_to(lm, red);
    __syncthreads();
    const float h = crit_half * (sqrtf((float)red_min(red)) * inv_c);
#pragma unroll
    for (int s = 0; s < NPL; ++s)
      if (s * B + tid < Nf) {
        float f[3];
        node_force(s, f);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[s][d] = vr[s][d] + h * (-f[d] * imr[s] - eta * vr[s][d]);
      }
    h_prev = h;
  }

  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  for (;;) {
    const float h2 = 2.0f * h_prev;
    // drift: u (hi, lo) += 2h v with Fast2Sum, exact when |u_hi| >= |increment|;
    // otherwise u is within a few increments of zero and the uncompensated
    // rounding is below eps * |increment|.
#pragma unroll
    for (int s = 0; s < NPL; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
        float pc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float y = ul[s][d] + h2 * vr[s][d];
          const float sh = uh[s][d] + y;
          ul[s][d] = y - (sh - uh[s][d]);
          uh[s][d] = sh;
          pc[d] = sh;
        }
        pos[r] = F3{pc[0], pc[1], pc[2]};
      }
    }
    if (ramp && tid < Nb) {  // waves without boundary nodes skip the amplitude
      const double alpha = ramp_alpha_step(step + 1, inv_S, a.amplitude);
      for (int b = tid; b < Nb; b += B) {
        float ub[3];
        bnd_step(bst + 6 * b, alpha, ub);
        pos[Nf + b] = F3{ub[0], ub[1], ub[2]};
      }
    }
    __syncthreads();  // positions published; every force read of the last step done
    const bool refresh = (step % freq) == 0;
    const float lm = fiber_phase(refresh);
    if (refresh) wave_min_to(lm, red);
    __syncthreads();  // fiber forces published
    float h = h_prev;
    if (refresh) h = crit_half * (sqrtf((float)red_min(red)) * inv_c);
    const float hk = h_prev + h;
    double rsum = 0.0;
#pragma unroll
    for (int s = 0; s < NPL; ++s)
      if (s * B + tid < Nf) {
        float f[3];
        node_force(s, f);
        if (refresh) rsum += (double)dot3f(f[0], f[1], f[2], f[0], f[1], f[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          vr[s][d] = vr[s][d] + hk * (-f[d] * imr[s] - eta * vr[s][d]);
      }
    if (refresh) {
      // residual in slots 1 of the scratch (slots 0 are read above)
      const double ws = wave_sum(rsum);
      if (lane == 0) red[(tid / kWave) * kRedStride + 1] = ws;
      __syncthreads();
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) sum += red[k * kRedStride + 1];
      residual = sqrt(sum);
    }
    h_prev = h;
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

  // finalize: low parts into the force records, boundary gathers in FP32
  // with exactly rounded reference vectors, stress in FP64
  F3 *Ph = pos;
  F3 *Pl = ffr;
  __syncthreads();  // every force and scratch read is done
#pragma unroll
  for (int s = 0; s < NPL; ++s) {
    const int r = s * B + tid;
    if (r < Nf) Pl[r] = F3{ul[s][0], ul[s][1], ul[s][2]};
  }
  const double alpha_end = ramp ? ramp_alpha(step, a.n_steps, a.amplitude) : 0.0;
  for (int b = tid; b < Nb; b += B) {
    double ub[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    const F3 hb{(float)ub[0], (float)ub[1], (float)ub[2]};
    Ph[Nf + b] = hb;
    Pl[Nf + b] = F3{(float)(ub[0] - (double)hb.x), (float)(ub[1] - (double)hb.y),
                    (float)(ub[2] - (double)hb.z)};
  }
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  for (int b = tid; b < Nb; b += B) {
    const int rb = Nf + b;
    const int64_t gr = nb + rb;
    double ub[3], xb[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    for (int d = 0; d < 3; ++d) xb[d] = X[3 * rb + d] + ub[d];
    const F3 hb = Ph[rb], lb = Pl[rb];
    const int k = rb / kWave;
    const int base = a.slice_base[R.slice_off + k] + (rb & (kWave - 1));
    const int dg = a.deg[gr];
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const int w = wv & 0xFFFFFF;
      const F3 qh = Ph[w], ql = Pl[w];
      const float ex = (float)(X[3 * w] - X[3 * rb]);
      const float ey = (float)(X[3 * w + 1] - X[3 * rb + 1]);
      const float ez = (float)(X[3 * w + 2] - X[3 * rb + 2]);
      const float gx = (qh.x - hb.x) + (ql.x - lb.x);
      const float gy = (qh.y - hb.y) + (ql.y - lb.y);
      const float gz = (qh.z - hb.z) + (ql.z - lb.z);
      const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
      const float sq = dot3f(ex + dx, ey + dy, ez + dz, gx, gy, gz);
      float t;
      if (LIN) {
        const double l0 = gl0[idx];
        t = (float)(EA_half / (l0 * l0 * l0)) * sq;
      } else {
        const float il = gil[idx];
        t = react32_tab(rtab, (uint32_t)wv >> 24, sq * (il * il), il);
      }
      f1 -= dx * t;
      f2 -= dy * t;
      f3 -= dz * t;
    }
    const double g1 = f1, g2 = f2, g3 = f3;
    st[0] += xb[0] * g1;
    st[1] += xb[1] * g2;
    st[2] += xb[2] * g3;
    st[3] += 0.5 * fma(xb[1], g3, xb[2] * g2);
    st[4] += 0.5 * fma(xb[0], g3, xb[2] * g1);
    st[5] += 0.5 * fma(xb[0], g2, xb[1] * g1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * gr + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d) a.u_out[3 * gr + d] = ub[d];
  }
#pragma unroll
  for (int s = 0; s < NPL; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      const int64_t g = nb + r;
      for (int d = 0; d < 3; ++d) {
        const double u = (double)uh[s][d] + (double)ul[s][d];
        if (flags & kWriteX) a.x_out[3 * g + d] = X[3 * r + d] + u;
        if (flags & kWriteU) a.u_out[3 * g + d] = u;
      }
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

}  // namespace kern

int chunk_occupancy(const Variant &v, size_t lds) {
  if (v.prec != MF_PREC_FP32C || !(v.flags & kEllInRegs) || (v.flags & kWaveRve)) return 0;
  int n = 0;
#define MF_REG_OCC(B, N, RMODE, FLAG)                                          \
  if (v.npt == N && v.block == B && (v.flags & (kLinearOnly | kMix2)) == FLAG) { \
    auto k = kern::solve32r<B, N, RMODE, 0>;                                   \
    if (lds > 64 * 1024)                                                       \
      (void)hipFuncSetAttribute((const void *)k,                               \
                                hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                (int)kMaxLds);                                 \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, B, lds) != hipSuccess) n = 0; \
    return n;                                                                  \
  }
#define MF_REG_OCC2(B, N)                                                      \
  MF_REG_OCC(B, N, kern::RM_LIN, kLinearOnly)                                  \
  MF_REG_OCC(B, N, kern::RM_MIX2, kMix2)
  MF_REG_OCC(256, 1, kern::RM_TABLE, 0)
  MF_REG_OCC(512, 1, kern::RM_TABLE, 0)
  MF_REG_OCC2(256, 1)
  MF_REG_OCC2(512, 1)
  MF_REG_OCC2(640, 1)
  MF_REG_OCC2(768, 1)
  MF_REG_OCC2(1024, 1)
  MF_REG_OCC(768, 2, kern::RM_LIN, kLinearOnly)
  MF_REG_OCC(256, 2, kern::RM_LIN, kLinearOnly)
  MF_REG_OCC(512, 3, kern::RM_LIN, kLinearOnly)
  MF_REG_OCC(512, 3, kern::RM_MIX2, kMix2)
  MF_REG_OCC(512, 4, kern::RM_LIN, kLinearOnly)
#undef MF_REG_OCC2
#undef MF_REG_OCC
  return 0;
}

hipError_t launch_solve32(const Variant &v, const SolveArgs &a, int n_blocks,
                  size_t lds, hipStream_t s) {
#define MF_VARIANT_LAUNCH(P, B, N, F)                                        \
  if constexpr (P == MF_PREC_FP32C)                                          \
    if (v.block == B && v.npt == N && v.flags == F) {                       \
    auto k = kern::solve32c<B, N, F>;                                          \
    if (lds > 64 * 1024)                                                     \
      (void)hipFuncSetAttribute((const void *)k,                             \
                                hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)kMaxLds);                               \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);               \
    return hipGetLastError();                                                \
  }
  MF_VARIANT_LIST(MF_VARIANT_LAUNCH)
#undef MF_VARIANT_LAUNCH
#define MF_REG_LAUNCH(B, N, RMODE, FLAG)                                       \
  if (v.npt == N && v.block == B && (v.flags & kEllInRegs) &&                  \
      !(v.flags & kWaveRve) && (v.flags & (kLinearOnly | kMix2)) == FLAG) {    \
    auto k = kern::solve32r<B, N, RMODE, 0>;                                   \
    if (lds > 64 * 1024)                                                       \
      (void)hipFuncSetAttribute((const void *)k,                               \
                                hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                (int)kMaxLds);                                 \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);                 \
    return hipGetLastError();                                                  \
  }
#define MF_REG_LAUNCH2(B, N)                                                   \
  MF_REG_LAUNCH(B, N, kern::RM_LIN, kLinearOnly)                               \
  MF_REG_LAUNCH(B, N, kern::RM_MIX2, kMix2)
  // one RVE per wave: W RVEs per workgroup, the group's solves in
  // ceil(n / W) workgroups, lds = W regions of the largest RVE's bytes
#define MF_WAVE_LAUNCH(W, N, RMODE, FLAG)                                      \
  if (v.npt == N && v.block == 64 * (W) && (v.flags & kWaveRve) &&             \
      (v.flags & (kLinearOnly | kMix2)) == FLAG) {                             \
    auto k = kern::solve32r<64 * (W), N, RMODE, W>;                            \
    SolveArgs aw = a;                                                          \
    aw.wave_lds = (int32_t)((lds + 15) & ~size_t(15));                        \
    const size_t lw = (size_t)(W) * (size_t)aw.wave_lds;                       \
    if (lw > kMaxLds) return hipErrorInvalidValue;                             \
    if (lw > 64 * 1024)                                                        \
      (void)hipFuncSetAttribute((const void *)k,                               \
                                hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                (int)kMaxLds);                                 \
    hipLaunchKernelGGL(k, dim3((n_blocks + (W) - 1) / (W)), dim3(64 * (W)), lw, s, aw); \
    return hipGetLastError();                                                  \
  }
#define MF_WAVE_LAUNCH2(W, N)                                                  \
  MF_WAVE_LAUNCH(W, N, kern::RM_LIN, kLinearOnly)                              \
  MF_WAVE_LAUNCH(W, N, kern::RM_MIX2, kMix2)
  MF_WAVE_LIST(MF_WAVE_LAUNCH2)
#undef MF_WAVE_LAUNCH2
#undef MF_WAVE_LAUNCH
  // only the instantiations that stay spill-free (the reaction-table mode
  // holds the table rows of every fiber: 256 and 512 threads only)
  MF_REG_LAUNCH(256, 1, kern::RM_TABLE, 0)
  MF_REG_LAUNCH(512, 1, kern::RM_TABLE, 0)
  MF_REG_LAUNCH2(256, 1)
  MF_REG_LAUNCH2(512, 1)
  MF_REG_LAUNCH2(640, 1)
  MF_REG_LAUNCH2(768, 1)
  MF_REG_LAUNCH2(1024, 1)
  // several nodes per thread
  MF_REG_LAUNCH(768, 2, kern::RM_LIN, kLinearOnly)
  MF_REG_LAUNCH(256, 2, kern::RM_LIN, kLinearOnly)
  MF_REG_LAUNCH(512, 3, kern::RM_LIN, kLinearOnly)
  MF_REG_LAUNCH(512, 3, kern::RM_MIX2, kMix2)
  MF_REG_LAUNCH(512, 4, kern::RM_LIN, kLinearOnly)
#undef MF_REG_LAUNCH2
#undef MF_REG_LAUNCH
#define MF_FPAR_LAUNCH(B, FPL, NPL, FL)                                       \
  if (v.block == B && v.npt == NPL &&                                         \
      v.flags == (kFiberPar | (FL) | ((FPL) << kFplShift))) {                 \
    auto k = kern::solve32f<B, FPL, NPL,                                      \
                            (FL) == kLinearOnly ? kern::RM_LIN                \
                            : (FL) == kMix2     ? kern::RM_MIX2               \
                                                : kern::RM_TABLE>;            \
    if (lds > 64 * 1024)                                                      \
      (void)hipFuncSetAttribute((const void *)k,                              \
                                hipFuncAttributeMaxDynamicSharedMemorySize,  \
                                (int)kMaxLds);                                \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);                \
    return hipGetLastError();                                                 \
  }
  MF_FIBER_PAR_LIST(MF_FPAR_LAUNCH)
#undef MF_FPAR_LAUNCH
#define MF_FSTR_LAUNCH(B, FPL, NPL, FL)                                       \
  if (v.block == B && v.npt == NPL &&                                         \
      v.flags == (kFiberStream | (FL) | ((FPL) << kFplShift))) {              \
    auto k = kern::solve32s<B, FPL, NPL,                                      \
                            (FL) == kLinearOnly ? kern::RM_LIN : kern::RM_MIX2>; \
    if (lds > 64 * 1024)                                                      \
      (void)hipFuncSetAttribute((const void *)k,                              \
                                hipFuncAttributeMaxDynamicSharedMemorySize,  \
                                (int)kMaxLds);                                \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);                \
    return hipGetLastError();                                                 \
  }
  MF_FIBER_STREAM_LIST(MF_FSTR_LAUNCH)
#undef MF_FSTR_LAUNCH
  return hipErrorInvalidValue;
}

}  // namespace mf
