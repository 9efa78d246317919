// FP32 + compensated explicit solve: FP32 stretch, force and update
// arithmetic, (hi, lo) FP32 positions advanced with TwoSum, FP64 stress.
#include "kernels_common.hpp"

namespace mf {
namespace kern {

// ============================================================================
// FP32 + compensated kernel
// ============================================================================
__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

template <int B, int NPT, int FL>
__global__ __launch_bounds__(B) void solve32c(SolveArgs a) {
  constexpr bool kEll = (FL & kEllInLds) != 0;
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
  float4 *pos = reinterpret_cast<float4 *>(smem);  // [2][Np]
  off += sizeof(float4) * 2 * Np;
  double *red = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * NW * kRedStride;
  double *bst = reinterpret_cast<double *>(smem + off);
  off += sizeof(double) * 6 * Nbp;
  ReactDev *rxs = reinterpret_cast<ReactDev *>(smem + off);
  off += align16(sizeof(ReactDev) * kMaxReactions);
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
  if (tid < R.n_react) rxs[tid] = a.reacts[R.react_off + tid];
  const bool lin_only =
      R.n_react == 1 && a.reacts[R.react_off].type == MF_REACTION_LINEAR;
  const float lin_EA_half = a.reacts[R.react_off].EA_half;

  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.solve_F[9 * (int64_t)sid + i];
  const bool ramp = a.mode == MF_MODE_RAMP;

  float xh[NPT][3], xl[NPT][3], vr[NPT][3], ar[NPT][3], imr[NPT];
  int degr[NPT], Dr[NPT], baser[NPT];
  auto init_node = [&](int r, double (&x)[3]) {
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
    float4 p;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double u = ramp ? up[k] : inc[k] + up[k];
      x[k] = a.X[3 * g + k] + u;
    }
    p.x = (float)x[0];
    p.y = (float)x[1];
    p.z = (float)x[2];
    p.w = 0.0f;
    pos[r] = p;
    pos[Np + r] = p;
    if (r >= Nf) {
      const int b = r - Nf;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        bst[k * Nbp + b] = x[k];
        bst[(3 + k) * Nbp + b] = ramp ? inc[k] : 0.0;
      }
    }
  };
  const float eta = (float)a.eta;
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    double x[3] = {0, 0, 0};
    if (r < N) init_node(r, x);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xh[s][d] = (float)x[d];
      xl[s][d] = (float)(x[d] - (double)xh[s][d]);
      vr[s][d] = 0.0f;
      ar[s][d] = 0.0f;
    }
    degr[s] = 0;
    Dr[s] = 0;
    baser[s] = 0;
    imr[s] = 0.0f;
    if (r < Nf) {
      imr[s] = (float)(1.0 / a.mass[nb + r]);
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

  const float inv_c = (float)(1.0 / R.c_sound);
  int cur = 0;
  auto node_force = [&](int s, int buf, float (&f)[3]) {
    const float4 *P = pos + buf * Np;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int D = __builtin_amdgcn_readfirstlane(Dr[s]);
    for (int j = 0; j < D; ++j) {
      if (j < degr[s]) {
        const int idx = baser[s] + j * kWave;
        const int32_t wv = Ew[idx];
        const float il = Eil[idx];
        const float4 q = P[wv & 0xFFFFFF];
        const float dx = (q.x - xh[s][0]) - xl[s][0];
        const float dy = (q.y - xh[s][1]) - xl[s][1];
        const float dz = (q.z - xh[s][2]) - xl[s][2];
        const float l2 = dx * dx + dy * dy + dz * dz;
        const float lam2 = l2 * (il * il);
        const float t = lin_only ? lin_EA_half * il * (lam2 - 1.0f)
                                 : react32_t(rxs[(uint32_t)wv >> 24], lam2, il);
        fx -= dx * t;
        fy -= dy * t;
        fz -= dz * t;
      }
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  };
  auto min_length2 = [&](int buf) {
    const float4 *P = pos + buf * Np;
    float lmin = 3.0e38f;
    const int32_t *fb = a.fib + 2 * R.fib_off;
    for (int e = tid; e < M; e += B) {
      const float4 p1 = P[fb[2 * e]], p2 = P[fb[2 * e + 1]];
      float x1 = p2.x - p1.x, x2 = p2.y - p1.y, x3 = p2.z - p1.z;
      lmin = fminf(lmin, x1 * x1 + x2 * x2 + x3 * x3);
    }
    return (double)lmin;
  };

  // initialization: forces at the start positions, first time step
  float fint[NPT][3];
  double resid_local = 0.0;
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    if (s * B + tid < Nf) {
      node_force(s, cur, fint[s]);
#pragma unroll
      for (int d = 0; d < 3; ++d) ar[s][d] = -fint[s][d] * imr[s] - eta * vr[s][d];
    }
  }
  double lmin2 = min_length2(cur);
  block_min_sum<B>(lmin2, resid_local, red);
  float dt_crit = sqrtf((float)lmin2) * inv_c;

  int step = 0;
  double residual = 10.0;
  int status = MF_STATUS_OK;
  const int freq = a.freq;
  const float crit_half = (float)(a.crit * 0.5);
  for (;;) {
    const float h = crit_half * dt_crit;
    const float h2 = 2.0f * h;
    const int nxt = cur ^ 1;
    float4 *Pn = pos + nxt * Np;
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      const int r = s * B + tid;
      if (r < Nf) {
        float4 p;
        float pc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float vl = vr[s][d] + h * ar[s][d];
          vr[s][d] = vl;
          // compensated position update: (hi, lo) += 2h*v
          float y = xl[s][d] + h2 * vl;
          float sh, e;
          two_sum(xh[s][d], y, sh, e);
          xh[s][d] = sh;
          xl[s][d] = e;
          pc[d] = sh;
        }
        p.x = pc[0];
        p.y = pc[1];
        p.z = pc[2];
        p.w = 0.0f;
        Pn[r] = p;
      }
    }
    if (ramp) {
      const double alpha = (double)(step + 1) / (double)a.n_steps;
      for (int b = tid; b < Nb; b += B) {
        float4 p;
        p.x = (float)(bst[b] + alpha * bst[3 * Nbp + b]);
        p.y = (float)(bst[Nbp + b] + alpha * bst[4 * Nbp + b]);
        p.z = (float)(bst[2 * Nbp + b] + alpha * bst[5 * Nbp + b]);
        p.w = 0.0f;
        Pn[Nf + b] = p;
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
        float rn = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          ar[s][d] = -fint[s][d] * imr[s] - eta * vr[s][d];
          rn += fint[s][d] * fint[s][d];
        }
        rsum += (double)rn;
      }
    }
    if (refresh) {
      double lm = min_length2(cur);
      block_min_sum<B>(lm, rsum, red);
      dt_crit = sqrtf((float)lm) * inv_c;
      residual = sqrt(rsum);
    }
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
      if (s * B + tid < Nf) {
#pragma unroll
        for (int d = 0; d < 3; ++d) vr[s][d] += h * ar[s][d];
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

  // finalize: publish the low parts so boundary gathers see full positions
  float4 *Ph = pos + cur * Np;
  float4 *Pl = pos + (cur ^ 1) * Np;
  __syncthreads();  // every read of the other buffer is done
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      float4 p;
      p.x = xl[s][0];
      p.y = xl[s][1];
      p.z = xl[s][2];
      p.w = 0.0f;
      Pl[r] = p;
    }
  }
  const double alpha_end = ramp ? 1.0 : 0.0;
  for (int b = tid; b < Nb; b += B) {
    float4 p;
    double xb[3];
    for (int d = 0; d < 3; ++d) xb[d] = bst[d * Nbp + b] + alpha_end * bst[(3 + d) * Nbp + b];
    p.x = (float)(xb[0] - (double)Ph[Nf + b].x);
    p.y = (float)(xb[1] - (double)Ph[Nf + b].y);
    p.z = (float)(xb[2] - (double)Ph[Nf + b].z);
    p.w = 0.0f;
    Pl[Nf + b] = p;
  }
  __syncthreads();
  double st[6] = {0, 0, 0, 0, 0, 0};
  const int flags = a.solve_flags[sid];
  for (int b = tid; b < Nb; b += B) {
    const int r = Nf + b;
    double xb[3];
    for (int d = 0; d < 3; ++d) xb[d] = bst[d * Nbp + b] + alpha_end * bst[(3 + d) * Nbp + b];
    const float4 hb = Ph[r], lb = Pl[r];
    const int k = r / kWave;
    const int base = a.slice_base[R.slice_off + k] + (r & (kWave - 1));
    const int dg = a.deg[nb + r];
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    for (int j = 0; j < dg; ++j) {
      const int idx = base + j * kWave;
      const int32_t wv = gw[idx];
      const float il = gil[idx];
      const int w = wv & 0xFFFFFF;
      const float4 qh = Ph[w], ql = Pl[w];
      const float dx = (qh.x - hb.x) + (ql.x - lb.x);
      const float dy = (qh.y - hb.y) + (ql.y - lb.y);
      const float dz = (qh.z - hb.z) + (ql.z - lb.z);
      const float l2 = dx * dx + dy * dy + dz * dz;
      const float lam2 = l2 * (il * il);
      const float t = lin_only ? lin_EA_half * il * (lam2 - 1.0f)
                               : react32_t(rxs[(uint32_t)wv >> 24], lam2, il);
      f1 -= dx * t;
      f2 -= dy * t;
      f3 -= dz * t;
    }
    const double g1 = f1, g2 = f2, g3 = f3;
    st[0] += xb[0] * g1;
    st[1] += xb[1] * g2;
    st[2] += xb[2] * g3;
    st[3] += 0.5 * (xb[1] * g3 + xb[2] * g2);
    st[4] += 0.5 * (xb[0] * g3 + xb[2] * g1);
    st[5] += 0.5 * (xb[0] * g2 + xb[1] * g1);
    if (flags & kWriteX)
      for (int d = 0; d < 3; ++d) a.x_out[3 * (nb + r) + d] = xb[d];
    if (flags & kWriteU)
      for (int d = 0; d < 3; ++d)
        a.u_out[3 * (nb + r) + d] = xb[d] - a.X[3 * (nb + r) + d];
  }
#pragma unroll
  for (int s = 0; s < NPT; ++s) {
    const int r = s * B + tid;
    if (r < Nf) {
      const int64_t g = nb + r;
      for (int d = 0; d < 3; ++d) {
        double x = (double)xh[s][d] + (double)xl[s][d];
        if (flags & kWriteX) a.x_out[3 * g + d] = x;
        if (flags & kWriteU) a.u_out[3 * g + d] = x - a.X[3 * g + d];
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
      bad = bad || !(v == v) || v - v != 0.0;
    }
    if (bad && status == MF_STATUS_OK) status = MF_STATUS_NAN;
    a.status[sid] = status;
    a.steps[sid] = step;
  }
}

}  // namespace kern

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
  return hipErrorInvalidValue;
}

}  // namespace mf
