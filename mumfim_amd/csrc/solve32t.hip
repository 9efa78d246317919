// FP32 tangent-linear solve: one workgroup integrates a ramp-mode solve
// (FP32 stretch/force/update arithmetic, compensated (hi, lo) positions, FP64
// stress, as solve32c) together with its derivative along one perturbation
// of the deformation gradient, dF = G (G = E_i F for tangent column i, E_i
// the unit Voigt strain).  Every quantity the step computes is
// differentiated where it is computed: the ramped boundary
// (d u_b = alpha G x_b), the drift and the merged half kicks, each fiber's
// t = f/l through dt/de2, the critical step through the shortest fiber
// (refreshed with it), and the boundary stress sum.  The derivative of the
// stress is exact up to FP32 rounding: no difference of two nearly equal
// solves is taken, so it holds at any strain scale.
//
// LDS: displacement hi parts and their derivatives (one buffer each, two
// barriers per step), fixed-point reference coordinates, boundary records
// (bnd_set, FP64) and dg in FP32; incidence lists stream from HBM.
#include "kernels_common.hpp"

namespace mf {
namespace kern {

// (min l^2, fiber index, d l^2) over the workgroup: the shortest fiber
// (lowest index on ties) and its derivative.
template <int B>
__device__ inline void block_argmin(float &v, int &idx, float &dv, double *red) {
  constexpr int NW = B / kWave;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float v2 = __shfl_xor(v, o, kWave);
    const int i2 = __shfl_xor(idx, o, kWave);
    const float d2 = __shfl_xor(dv, o, kWave);
    if (v2 < v || (v2 == v && i2 < idx)) {
      v = v2;
      idx = i2;
      dv = d2;
    }
  }
  float *r = reinterpret_cast<float *>(red);
  if (lane == 0) {
    r[w * 2 * kRedStride + 0] = v;
    reinterpret_cast<int *>(r)[w * 2 * kRedStride + 1] = idx;
    r[w * 2 * kRedStride + 2] = dv;
  }
  __syncthreads();
  v = r[0];
  idx = reinterpret_cast<int *>(r)[1];
  dv = r[2];
#pragma unroll
  for (int k = 1; k < NW; ++k) {
    const float vk = r[k * 2 * kRedStride + 0];
    const int ik = reinterpret_cast<int *>(r)[k * 2 * kRedStride + 1];
    if (vk < v || (vk == v && ik < idx)) {
      v = vk;
      idx = ik;
      dv = r[k * 2 * kRedStride + 2];
    }
  }
}

// Sums NV (<= kRedStride) values over the workgroup into thread 0 only:
// waves reduce by shuffles, thread 0 adds the wave partials in wave order
// with a rolled loop (an unrolled one would hold every partial in registers).
template <int B, int NV>
__device__ inline void block_sum_lane0(double (&v)[NV], double *red) {
  constexpr int NW = B / kWave;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double s = wave_sum(v[i]);
    if (lane == 0) red[w * kRedStride + i] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = red[i];
#pragma unroll 1
    for (int k = 1; k < NW; ++k)
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] += red[k * kRedStride + i];
  }
}

template <int B, int NPT>
__global__ __launch_bounds__(B) void solve32t(SolveArgs a) {
  constexpr int NW = B / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sid = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  if (!a.solve_G || !a.dsigma) {  // launched outside a tangent-linear set
    if (tid == 0) {
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

  size_t off = 0;
  float4 *pos = reinterpret_cast<float4 *>(smem);  // [Np] displacement hi
  off += sizeof(float4) * Np;
  float4 *dpos = reinterpret_cast<float4 *>(smem + off);  // [Np] its derivative
  off += sizeof(float4) * Np;
  int4 *Xs = reinterpret_cast<int4 *>(smem + off);  // [Np] reference X, fixed point
  off += sizeof(int4) * Np;
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);  // [Nbp][6] u_b0, g
  off += sizeof(double) * 6 * Nbp;
  float4 *dbst = reinterpret_cast<float4 *>(smem + off);  // [Nbp] dg = G x_b
  off += sizeof(float4) * Nbp;
  float4 *rtab = reinterpret_cast<float4 *>(smem + off);  // packed reaction table

  const int32_t *gw = a.ell_w + R.ell_off;
  const float *gil = a.ell_il + R.ell_off;
  if (tid < R.n_react) pack_react_table(a.reacts[R.react_off + tid], rtab + kReactTabStride * tid);
  int nl_rid = -1;
  float ea_half_lin = 0.0f;
  const bool mix2 = mix2_table(a.reacts + R.react_off, R.n_react, nl_rid, ea_half_lin);
  const ReactDev nlq = a.reacts[R.react_off + (nl_rid < 0 ? 0 : nl_rid)];
  const bool lin_only =
      R.n_react == 1 && a.reacts[R.react_off].type == MF_REACTION_LINEAR;
  const float lin_EA_half = a.reacts[R.react_off].EA_half;

  double F[9], G[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F[i] = a.solve_F[9 * (int64_t)sid + i];
    G[i] = a.solve_G[9 * (int64_t)sid + i];
  }
  const double inv_S = 1.0 / a.n_steps;

  float uh[NPT][3], ul[NPT][3], vr[NPT][3], du[NPT][3], dv[NPT][3], imr[NPT];
  // slice width and base (wave-uniform); Dn: with one linear and one
  // nonlinear reaction, entries [0, Dn) hold the wave's nonlinear fibers
  // (nodes list them first) and later entries are linear on every lane, so
  // the wave skips the toe law and its derivative there
  int Dr[NPT], sbase[NPT], Dn[NPT];
  // ramp mode: u kept, boundary increment (F - I) x and its derivative G x
  // set aside; derivatives start at zero (the start state does not depend
  // on the perturbation)
  auto init_node = [&](int r, double (&u)[3]) {
    const int64_t g = nb + r;
    double xp[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = a.x_in[3 * g + d];
      u[d] = a.u_in[3 * g + d];
    }
    pos[r] = make_float4((float)u[0], (float)u[1], (float)u[2], 0.0f);
    dpos[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    Xs[r] = make_int4((int)rint(a.X[3 * g] * R.xq_scale),
                      (int)rint(a.X[3 * g + 1] * R.xq_scale),
                      (int)rint(a.X[3 * g + 2] * R.xq_scale), 0);
    if (r >= Nf) {
      const int b = r - Nf;
      double inc[3];
      float dg[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        inc[k] = (F[3 * k + 0] - (k == 0 ? 1.0 : 0.0)) * xp[0] +
                 (F[3 * k + 1] - (k == 1 ? 1.0 : 0.0)) * xp[1] +
                 (F[3 * k + 2] - (k == 2 ? 1.0 : 0.0)) * xp[2];
        dg[k] = (float)(G[3 * k + 0] * xp[0] + G[3 * k + 1] * xp[1] + G[3 * k + 2] * xp[2]);
      }
      bnd_set(bst + 6 * b, u, inc);
      dbst[b] = make_float4(dg[0], dg[1], dg[2], 0.0f);
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
      vr[s][d] = du[s][d] = dv[s][d] = 0.0f;
    }
    imr[s] = 0.0f;
    if (r < Nf) imr[s] = (float)(1.0 / a.mass[nb + r]);
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
  // t = f/l and dt/de2 of one incidence
  auto react_d = [&](int32_t wv, float e2, float il, float &dt) {
    const unsigned rid = (uint32_t)wv >> 24;
    if (lin_only) {
      dt = lin_EA_half * il;
      return dt * e2;
    }
    return mix2 ? react32_mix2_d(nlq, ea_half_lin, (int)rid == nl_rid, e2, il, dt)
                : react32_tab_d(rtab, rid, e2, il, dt);
  };
  // internal force of a slot's node and its derivative (ELL padding names
  // the node itself with 1/l0 = 0: zero force, zero derivative)
  auto node_force = [&](int s, float (&f)[3], float (&df)[3]) {
    const int4 Xv = Xs[s * B + tid];
    const int base = sbase[s] + lane;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f, gx_ = 0.0f, gy_ = 0.0f, gz_ = 0.0f;
    const int D = Dr[s];
    if (D > 0) {
      int32_t wv_n = gw[base];
      float il_n = gil[base];
      for (int j = 0; j < D; ++j) {
        const int32_t wv = wv_n;
        const float il = il_n;
        if (j + 1 < D) {
          wv_n = gw[base + (j + 1) * kWave];
          il_n = gil[base + (j + 1) * kWave];
        }
        const int w = wv & 0xFFFFFF;
        const float4 q = ld16(pos + w);
        const float4 dq = ld16(dpos + w);
        const int4 xq = ld16(Xs + w);
        const float ex = (float)(xq.x - Xv.x) * xq_inv;
        const float ey = (float)(xq.y - Xv.y) * xq_inv;
        const float ez = (float)(xq.z - Xv.z) * xq_inv;
        const float gx = (q.x - uh[s][0]) - ul[s][0];
        const float gy = (q.y - uh[s][1]) - ul[s][1];
        const float gz = (q.z - uh[s][2]) - ul[s][2];
        const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
        const float sl = (ex + dx) * gx + (ey + dy) * gy + (ez + dz) * gz;
        const float il2 = il * il;
        float dt, t;
        if (mix2 && j >= Dn[s]) {  // linear on every lane of the wave
          dt = ea_half_lin * il;
          t = dt * (sl * il2);
        } else {
          t = react_d(wv, sl * il2, il, dt);
        }
        // d(l^2 - l0^2) = 2 d . dd
        const float ddx = dq.x - du[s][0], ddy = dq.y - du[s][1], ddz = dq.z - du[s][2];
        const float de2 = 2.0f * (dx * ddx + dy * ddy + dz * ddz) * il2;
        const float dtt = dt * de2;
        fx -= dx * t;
        fy -= dy * t;
        fz -= dz * t;
        gx_ -= ddx * t + dx * dtt;
        gy_ -= ddy * t + dy * dtt;
        gz_ -= ddz * t + dz * dtt;
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
    df[0] = gx_;
    df[1] = gy_;
    df[2] = gz_;
  };
  // shortest fiber's l^2 and its derivative (every dt refresh)
  auto min_length2 = [&](float &dl2) {
    float lmin = 3.0e38f, dmin = 0.0f;
    int imin = 0x7FFFFFFF;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += B) {
      const int n1 = fb[2 * e], n2 = fb[2 * e + 1];
      const float4 p1 = pos[n1], p2 = pos[n2];
      const float4 q1 = dpos[n1], q2 = dpos[n2];
      const int4 X1 = Xs[n1], X2 = Xs[n2];
      const float x1 = (float)(X2.x - X1.x) * xq_inv + (p2.x - p1.x);
      const float x2 = (float)(X2.y - X1.y) * xq_inv + (p2.y - p1.y);
      const float x3 = (float)(X2.z - X1.z) * xq_inv + (p2.z - p1.z);
      const float l2 = x1 * x1 + x2 * x2 + x3 * x3;
      if (l2 < lmin) {  // fibers visited in increasing index per thread
        lmin = l2;
        imin = e;
        dmin = 2.0f * (x1 * (q2.x - q1.x) + x2 * (q2.y - q1.y) + x3 * (q2.z - q1.z));
      }
    }
    block_argmin<B>(lmin, imin, dmin, red);
    dl2 = dmin;
    return lmin;
  };

  const float crit_half = (float)(a.crit * 0.5);
  // h = crit/2 sqrt(l2) / c and dh = crit/2 dl2 / (2 sqrt(l2) c)
  auto step_size = [&](float l2, float dl2, float &dh) {
    const float l = sqrtf(l2);
    dh = crit_half * (0.5f * dl2 / l) * inv_c;
    return crit_half * (l * inv_c);
  };
  float h_prev = 0.0f, dh_prev = 0.0f;
  {
    float f[NPT][3], df[NPT][3];
#pragma unroll
    for (int s = 0; s < NPT; ++s)
      if (s * B + tid < Nf) node_force(s, f[s], df[s]);
    float dl2;
    const float l2 = min_length2(dl2);
    float dh;
    const float h = step_size(l2, dl2, dh);
#pragma unroll
    for (int s = 0; s < NPT; ++s)
      if (s * B + tid < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float acc = -f[s][d] * imr[s] - eta * vr[s][d];
          const float dacc = -df[s][d] * imr[s] - eta * dv[s][d];
          vr[s][d] = vr[s][d] + h * acc;
          dv[s][d] = dv[s][d] + dh * acc + h * dacc;
        }
      }
    h_prev = h;
    dh_prev = dh;
  }

  int step = 0;
  double residual = 10.0;
  const int freq = a.freq;
  for (;;) {
    const float h2 = 2.0f * h_prev, dh2 = 2.0f * dh_prev;
    __syncthreads();  // every read of the last step's positions done
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
        float pc[3], pd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float y = ul[s][d] + h2 * vr[s][d];
          const float sh = uh[s][d] + y;
          ul[s][d] = y - (sh - uh[s][d]);
          uh[s][d] = sh;
          pc[d] = sh;
          du[s][d] = du[s][d] + (dh2 * vr[s][d] + h2 * dv[s][d]);
          pd[d] = du[s][d];
        }
        pos[r] = make_float4(pc[0], pc[1], pc[2], 0.0f);
        dpos[r] = make_float4(pd[0], pd[1], pd[2], 0.0f);
      }
    }
    if (tid < Nb) {
      const double alpha = ramp_alpha_step(step + 1, inv_S, a.amplitude);
      const float af = (float)alpha;
      for (int b = tid; b < Nb; b += B) {
        float ub[3];
        bnd_step(bst + 6 * b, alpha, ub);
        pos[Nf + b] = make_float4(ub[0], ub[1], ub[2], 0.0f);
        const float4 g = dbst[b];
        dpos[Nf + b] = make_float4(af * g.x, af * g.y, af * g.z, 0.0f);
      }
    }
    __syncthreads();
    const bool refresh = (step % freq) == 0;
    float h = h_prev, dh = dh_prev;
    if (refresh) {
      float dl2;
      const float l2 = min_length2(dl2);
      h = step_size(l2, dl2, dh);
    }
    const float hk = h_prev + h, dhk = dh_prev + dh;
    double rsum = 0.0;
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
        float f[3], df[3];
        node_force(s, f, df);
        if (refresh) rsum += (double)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float acc = -f[d] * imr[s] - eta * vr[s][d];
          const float dacc = -df[d] * imr[s] - eta * dv[s][d];
          vr[s][d] = vr[s][d] + hk * acc;
          dv[s][d] = dv[s][d] + (dhk * acc + hk * dacc);
        }
      }
    }
    if (refresh) {
      double none = 0.0;
      block_min_sum<B>(none, rsum, red + 4);
      residual = sqrt(rsum);
    }
    h_prev = h;
    dh_prev = dh;
    ++step;
    if (step >= a.n_steps) break;
  }
  int status = residual != residual ? MF_STATUS_NAN : MF_STATUS_OK;

  const double alpha_end = ramp_alpha(step, a.n_steps, a.amplitude);
  const int flags = a.solve_flags[sid];
  __syncthreads();  // every read of the step's positions done
  for (int b = tid; b < Nb; b += B) {
    double ub[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    pos[Nf + b] = make_float4((float)ub[0], (float)ub[1], (float)ub[2], 0.0f);
    const float4 g = dbst[b];
    const float af = (float)alpha_end;
    dpos[Nf + b] = make_float4(af * g.x, af * g.y, af * g.z, 0.0f);
  }
  __syncthreads();
  // boundary force derivatives from the hi parts and the derivatives:
  // d sigma = s/V sum_b sym(dx_b (x) f_b + x_b (x) df_b)
  double ds[6] = {0, 0, 0, 0, 0, 0};
  auto bgather = [&](int r, const float4 hb, const float4 lb, const float4 *Pl, bool with_d,
                     float (&f)[3], float (&df)[3]) {
    const int4 Xb = Xs[r];
    const float4 db = dpos[r];
    const int k = r / kWave;
    const int base = a.slice_base[R.slice_off + k] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const float il = gil[idx];
      const int w = wv & 0xFFFFFF;
      const float4 qh = pos[w];
      const float4 ql = Pl ? Pl[w] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const int4 xq = Xs[w];
      const float ex = (float)(xq.x - Xb.x) * xq_inv;
      const float ey = (float)(xq.y - Xb.y) * xq_inv;
      const float ez = (float)(xq.z - Xb.z) * xq_inv;
      const float gx = (qh.x - hb.x) + (ql.x - lb.x);
      const float gy = (qh.y - hb.y) + (ql.y - lb.y);
      const float gz = (qh.z - hb.z) + (ql.z - lb.z);
      const float dx = ex + gx, dy = ey + gy, dz = ez + gz;
      const float sl = (ex + dx) * gx + (ey + dy) * gy + (ez + dz) * gz;
      const float il2 = il * il;
      float dt;
      const float t = react_d(wv, sl * il2, il, dt);
      f1 -= dx * t;
      f2 -= dy * t;
      f3 -= dz * t;
      if (with_d) {
        const float4 dq = dpos[w];
        const float ddx = dq.x - db.x, ddy = dq.y - db.y, ddz = dq.z - db.z;
        const float dtt = dt * 2.0f * (dx * ddx + dy * ddy + dz * ddz) * il2;
        g1 -= ddx * t + dx * dtt;
        g2 -= ddy * t + dy * dtt;
        g3 -= ddz * t + dz * dtt;
      }
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
    df[0] = g1;
    df[1] = g2;
    df[2] = g3;
  };
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    double ub[3], xb[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    for (int d = 0; d < 3; ++d) xb[d] = a.X[3 * (nb + r) + d] + ub[d];
    float f[3], df[3];
    bgather(r, pos[r], make_float4(0.0f, 0.0f, 0.0f, 0.0f), nullptr, true, f, df);
    const float4 db = dpos[r];
    const double dxb[3] = {db.x, db.y, db.z};
    const double fb[3] = {f[0], f[1], f[2]}, dfb[3] = {df[0], df[1], df[2]};
    ds[0] += dxb[0] * fb[0] + xb[0] * dfb[0];
    ds[1] += dxb[1] * fb[1] + xb[1] * dfb[1];
    ds[2] += dxb[2] * fb[2] + xb[2] * dfb[2];
    ds[3] += 0.5 * (dxb[1] * fb[2] + xb[1] * dfb[2] + dxb[2] * fb[1] + xb[2] * dfb[1]);
    ds[4] += 0.5 * (dxb[0] * fb[2] + xb[0] * dfb[2] + dxb[2] * fb[0] + xb[2] * dfb[0]);
    ds[5] += 0.5 * (dxb[0] * fb[1] + xb[0] * dfb[1] + dxb[1] * fb[0] + xb[1] * dfb[0]);
  }
  __syncthreads();  // the derivative buffer is free: it takes the low parts
  float4 *Pl = dpos;
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    if (r < Nf) Pl[r] = make_float4(ul[s][0], ul[s][1], ul[s][2], 0.0f);
  }
  for (int b = tid; b < Nb; b += B) {
    double ub[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    const float4 hb = pos[Nf + b];
    Pl[Nf + b] = make_float4((float)(ub[0] - (double)hb.x), (float)(ub[1] - (double)hb.y),
                             (float)(ub[2] - (double)hb.z), 0.0f);
  }
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    const int64_t gr = nb + r;
    double ub[3], xb[3];
    bnd_disp(bst + 6 * b, alpha_end, ub);
    for (int d = 0; d < 3; ++d) xb[d] = a.X[3 * gr + d] + ub[d];
    float f[3], df[3];
    bgather(r, pos[r], Pl[r], Pl, false, f, df);
    const double g1 = f[0], g2 = f[1], g3 = f[2];
    st[0] += xb[0] * g1;
    st[1] += xb[1] * g2;
    st[2] += xb[2] * g3;
    st[3] += 0.5 * (xb[1] * g3 + xb[2] * g2);
    st[4] += 0.5 * (xb[0] * g3 + xb[2] * g1);
    st[5] += 0.5 * (xb[0] * g2 + xb[1] * g1);
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
  block_sum_lane0<B, 6>(st, red);
  __syncthreads();  // thread 0 has read red
  block_sum_lane0<B, 6>(ds, red);
  if (tid == 0) {
    const double sv = R.scale / a.V[rv];
    bool bad = false;
    for (int i = 0; i < 6; ++i) {
      const double v = st[i] * sv, w = ds[i] * sv;
      a.sigma[6 * (int64_t)sid + i] = v;
      a.dsigma[6 * (int64_t)sid + i] = w;
      bad = bad || !isfinite(v) || !isfinite(w);
    }
    if (bad && status == MF_STATUS_OK) status = MF_STATUS_NAN;
    a.status[sid] = status;
    a.steps[sid] = step;
  }
}

}  // namespace kern

// Tangent-linear variants: (threads, free-node slots per thread).
hipError_t launch_solve32t(const Variant &v, const SolveArgs &a, int n_blocks, size_t lds,
                           hipStream_t s) {
#define MF_TL_LAUNCH(B, N)                                                   \
  if (v.block == B && v.npt == N) {                                          \
    auto k = kern::solve32t<B, N>;                                           \
    if (lds > 64 * 1024)                                                     \
      (void)hipFuncSetAttribute((const void *)k,                             \
                                hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)kMaxLds);                               \
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(B), lds, s, a);               \
    return hipGetLastError();                                                \
  }
  MF_TL_LIST(MF_TL_LAUNCH)
#undef MF_TL_LAUNCH
  return hipErrorInvalidValue;
}

}  // namespace mf
