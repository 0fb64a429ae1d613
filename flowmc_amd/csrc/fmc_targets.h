// Built-in log-densities: value and hand-written gradient, evaluated by the
// lanes of one chain group.
//
// Lane layout: a chain is held by `lpc` consecutive lanes (lane r = 0..lpc-1
// of the group whose first lane is `gbase`); lane r holds NV float4 chunks,
// element e of chunk v being dimension  j = (v*lpc + r)*4 + e  (j >= d is
// padding and must hold 0). Every function returns the full log-density in
// every lane of the group and writes the lane's gradient chunks.
//
// These replace jax.value_and_grad(logpdf) of the user target at
// src/flowMC/resource/kernel/MALA.py:59 and jax.grad(potential) at
// src/flowMC/resource/kernel/HMC.py:75-77 for the targets listed in
// include/flowmc.h.
#pragma once

#include "fmc_common.h"

namespace fmc {

struct TargetDev {
  int kind;
  int d;
  int K;           // modes (GMM)
  const float* p;  // parameter block (LDS copy or global)
};

__device__ __forceinline__ int dim_of(int v, int e, int r, int lpc) { return (v * lpc + r) * 4 + e; }

template <int NV>
__device__ __forceinline__ float tgt_gauss_diag(const TargetDev& t, const float4 (&x)[NV],
                                                float4 (&g)[NV], int r, int lpc) {
  const float* mean = t.p;
  const float* prec = t.p + t.d;
  float acc = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int j = dim_of(v, e, r, lpc);
      float gj = 0.f;
      if (j < t.d) {
        float y = f4c(x[v], e) - mean[j];
        float pr = prec[j];
        gj = -pr * y;
        acc += 0.5f * gj * y;
      }
      f4(g[v], e) = gj;
    }
  return group_sum(acc, lpc);
}

template <int NV>
__device__ __forceinline__ float tgt_gauss_dense(const TargetDev& t, const float4 (&x)[NV],
                                                 float4 (&g)[NV], int r, int lpc, int gbase) {
  const float* mean = t.p;
  const float* P = t.p + t.d;  // symmetric [d,d]
  const int d = t.d;
  float4 y[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int j = dim_of(v, e, r, lpc);
      f4(y[v], e) = (j < d) ? f4c(x[v], e) - mean[j] : 0.f;
      f4(g[v], e) = 0.f;
    }
  // g_i = -sum_j P_ij y_j; y_j gathered from the lane that owns it.
  for (int s = 0; s < lpc; ++s) {
#pragma unroll
    for (int v2 = 0; v2 < NV; ++v2)
#pragma unroll
      for (int e2 = 0; e2 < 4; ++e2) {
        float ys = __shfl(f4c(y[v2], e2), gbase + s, 64);
        int j2 = dim_of(v2, e2, s, lpc);
        if (j2 < d) {
#pragma unroll
          for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              int j = dim_of(v, e, r, lpc);
              if (j < d) f4(g[v], e) -= P[(int64_t)j * d + j2] * ys;
            }
        }
      }
  }
  float acc = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc += 0.5f * f4c(y[v], e) * f4c(g[v], e);
  return group_sum(acc, lpc);
}

// log p = -scale * sum_{i<d-1} [100 (x_{i+1} - x_i^2)^2 + (1 - x_i)^2]
template <int NV>
__device__ __forceinline__ float tgt_rosenbrock(const TargetDev& t, const float4 (&x)[NV],
                                                float4 (&g)[NV], int r, int lpc, int gbase) {
  const float scale = t.p[0];
  const int d = t.d;
  float acc = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    // x_{j+1} for element 3 and x_{j-1} for element 0 live in neighbouring lanes.
    float nx_same = __shfl(x[v].x, gbase + ((r + 1) & (lpc - 1)), 64);
    float nx_next = group_bcast0(x[(v + 1 < NV) ? v + 1 : v].x, lpc);
    float pw_same = __shfl(x[v].w, gbase + ((r - 1) & (lpc - 1)), 64);
    float pw_prev = __shfl(x[(v > 0) ? v - 1 : 0].w, gbase + lpc - 1, 64);
    float nxt3 = (r < lpc - 1) ? nx_same : nx_next;
    float prv0 = (r > 0) ? pw_same : pw_prev;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int j = dim_of(v, e, r, lpc);
      float xi = f4c(x[v], e);
      float xn = (e < 3) ? f4c(x[v], e + 1) : nxt3;
      float xp = (e > 0) ? f4c(x[v], e - 1) : prv0;
      float gi = 0.f;
      if (j < d - 1) {
        float tt = xn - xi * xi;
        float om = 1.f - xi;
        acc += 100.f * tt * tt + om * om;
        gi += -400.f * xi * tt - 2.f * om;
      }
      if (j > 0 && j < d) gi += 200.f * (xi - xp * xp);
      f4(g[v], e) = -scale * gi;
    }
  }
  return -scale * group_sum(acc, lpc);
}

// Neal's funnel: x_0 = v ~ N(0, 3^2), x_i | v ~ N(0, e^v) (unnormalised).
// Padding elements (j >= d) hold 0 on entry, so they add nothing to the sum
// and get a zero gradient without a per-element index test; only element 0
// of the group's first lane (v itself) is special. exp(-v) is capped at
// FLT_MAX so that 0 * exp(-v) stays 0 on padding when e^-v overflows.
template <int NV>
__device__ __forceinline__ float tgt_funnel(const TargetDev& t, const float4 (&x)[NV],
                                            float4 (&g)[NV], int r, int lpc, int gbase) {
  const int d = t.d;
  const float v0 = group_bcast0(x[0].x, lpc);
  const float x00 = (r == 0) ? 0.f : x[0].x;
  float s = x00 * x00;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = (v == 0) ? 1 : 0; e < 4; ++e) s += f4c(x[v], e) * f4c(x[v], e);
  s = group_sum(s, lpc);
  const float em = fminf(expf(-v0), 3.402823466e38f);
  const float nm1 = (float)(d - 1);
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) f4(g[v], e) = -f4c(x[v], e) * em;
  const float g0 = -v0 * (1.f / 9.f) - 0.5f * nm1 + 0.5f * em * s;
  if (r == 0) g[0].x = g0;
  return -v0 * v0 * (1.f / 18.f) - 0.5f * nm1 * v0 - 0.5f * em * s;
}

// Diagonal Gaussian mixture: log p = logsumexp_k [ c_k - 1/2 sum_j ((x_j - mu_kj) / s_kj)^2 ]
// with c_k = log w_k - sum_j log s_kj. Online softmax over modes.
template <int NV>
__device__ __forceinline__ float tgt_gmm_diag(const TargetDev& t, const float4 (&x)[NV],
                                              float4 (&g)[NV], int r, int lpc) {
  const int d = t.d, K = t.K;
  const float* mu = t.p;
  const float* is = t.p + (int64_t)K * d;
  const float* c = t.p + 2 * (int64_t)K * d;
  float m = -INFINITY, sw = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v) g[v] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int k = 0; k < K; ++k) {
    float4 gk[NV];
    float q = 0.f;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int j = dim_of(v, e, r, lpc);
        float gj = 0.f;
        if (j < d) {
          float isg = is[(int64_t)k * d + j];
          float z = (f4c(x[v], e) - mu[(int64_t)k * d + j]) * isg;
          q += z * z;
          gj = -z * isg;
        }
        f4(gk[v], e) = gj;
      }
    q = c[k] - 0.5f * group_sum(q, lpc);
    if (q > m) {  // group-uniform branch: q is bitwise equal across the group
      float sc = expf(m - q);
      sw = sw * sc + 1.f;
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) f4(g[v], e) = f4c(g[v], e) * sc + f4c(gk[v], e);
      m = q;
    } else {
      float w = expf(q - m);
      sw += w;
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) f4(g[v], e) += w * f4c(gk[v], e);
    }
  }
  float inv = 1.f / sw;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) f4(g[v], e) *= inv;
  return m + logf(sw);
}

// Dual moon (docs/tutorials/dualmoon.ipynb, test/integration/test_MALA.py:11-20):
// log p = -50 (|x - center| - 2)^2 + LSE_a(-(x0 + a)^2 / (2*0.8^2)) + LSE_a(-(x1 + a)^2 / (2*0.6^2)),
// a in {-3, +3}. Requires d >= 2.
template <int NV>
__device__ __forceinline__ float tgt_dual_moon(const TargetDev& t, const float4 (&x)[NV],
                                               float4 (&g)[NV], int r, int lpc, int gbase) {
  const int d = t.d;
  const float* ctr = t.p;
  float n2 = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int j = dim_of(v, e, r, lpc);
      if (j < d) {
        float y = f4c(x[v], e) - ctr[j];
        n2 += y * y;
      }
    }
  n2 = group_sum(n2, lpc);
  float n = sqrtf(n2);
  float x0 = group_bcast0(x[0].x, lpc);
  float x1 = group_bcast0(x[0].y, lpc);
  // logsumexp pieces
  float a0 = -0.5f * ((x0 - 3.f) / 0.8f) * ((x0 - 3.f) / 0.8f);
  float b0 = -0.5f * ((x0 + 3.f) / 0.8f) * ((x0 + 3.f) / 0.8f);
  float m0 = fmaxf(a0, b0);
  float ea0 = expf(a0 - m0), eb0 = expf(b0 - m0);
  float l0 = m0 + logf(ea0 + eb0);
  float a1 = -0.5f * ((x1 - 3.f) / 0.6f) * ((x1 - 3.f) / 0.6f);
  float b1 = -0.5f * ((x1 + 3.f) / 0.6f) * ((x1 + 3.f) / 0.6f);
  float m1 = fmaxf(a1, b1);
  float ea1 = expf(a1 - m1), eb1 = expf(b1 - m1);
  float l1 = m1 + logf(ea1 + eb1);
  float t1 = (n - 2.f) / 0.1f;
  float radial = -100.f * (n - 2.f) / n;  // d/dx_j of -50 (n-2)^2 is radial * y_j
  float g0 = (ea0 * (-(x0 - 3.f)) + eb0 * (-(x0 + 3.f))) / ((ea0 + eb0) * 0.64f);
  float g1 = (ea1 * (-(x1 - 3.f)) + eb1 * (-(x1 + 3.f))) / ((ea1 + eb1) * 0.36f);
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int j = dim_of(v, e, r, lpc);
      float gj = 0.f;
      if (j < d) {
        gj = radial * (f4c(x[v], e) - ctr[j]);
        if (j == 0) gj += g0;
        if (j == 1) gj += g1;
      }
      f4(g[v], e) = gj;
    }
  return -(0.5f * t1 * t1 - l0 - l1);
}

// TK >= 0: the kind is fixed at compile time (one target per kernel);
// TK = -1: dispatch on t.kind at run time.
template <int NV, int TK = -1>
__device__ __forceinline__ float target_vg(const TargetDev& t, const float4 (&x)[NV],
                                           float4 (&g)[NV], int r, int lpc, int gbase) {
  switch (TK >= 0 ? TK : t.kind) {
    case FMC_TARGET_GAUSS_DIAG:
      return tgt_gauss_diag<NV>(t, x, g, r, lpc);
    case FMC_TARGET_GAUSS_DENSE:
      return tgt_gauss_dense<NV>(t, x, g, r, lpc, gbase);
    case FMC_TARGET_ROSENBROCK:
      return tgt_rosenbrock<NV>(t, x, g, r, lpc, gbase);
    case FMC_TARGET_FUNNEL:
      return tgt_funnel<NV>(t, x, g, r, lpc, gbase);
    case FMC_TARGET_GMM_DIAG:
      return tgt_gmm_diag<NV>(t, x, g, r, lpc);
    case FMC_TARGET_DUAL_MOON:
    default:
      return tgt_dual_moon<NV>(t, x, g, r, lpc, gbase);
  }
}

// Number of floats of the parameter block of each kind (host check).
inline int64_t target_param_count(int kind, int d, int K) {
  switch (kind) {
    case FMC_TARGET_GAUSS_DIAG: return 2LL * d;
    case FMC_TARGET_GAUSS_DENSE: return (int64_t)d + (int64_t)d * d;
    case FMC_TARGET_ROSENBROCK: return 1;
    case FMC_TARGET_FUNNEL: return 0;
    case FMC_TARGET_GMM_DIAG: return 2LL * K * d + K;
    case FMC_TARGET_DUAL_MOON: return d;
    default: return -1;
  }
}

}  // namespace fmc
