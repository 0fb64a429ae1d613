// Fused chain steps shared by the chain kernels (fmc_chains.hip) and the
// batched uneven-dimension HMC kernel (fmc_hmc_batched.hip): buffer I/O of a
// chain group, Philox draws, and the MALA / GRW / HMC step loops of one chain
// held by `lpc` lanes of a wavefront (lpc a power of two <= 64, groups
// aligned inside the wave); each lane keeps NV float4 chunks of position /
// gradient in registers for the whole launch.
#pragma once

#include "fmc_common.h"
#include "fmc_targets.h"

namespace fmc {

struct ChainIO {
  float* x;
  float* logp;
  const float* metric;
  const float* noise;
  const float* log_u;
  float* out_pos;
  float* out_logp;
  float* out_acc;
  int64_t C;
  int64_t buf_len;
  int64_t cursor;
  uint32_t key0, key1;
  const float* temps;  // tempered rows: row c targets log p / temps[c % n_temps] (NULL: untempered)
  int n_temps;
  // HMC with a dense condition matrix M: chol = the lower-triangular
  // Cholesky factor of inv(M) ([d, d] row-major), and metric = the kinetic
  // weights colsum(M) (NULL: diagonal or unit metric)
  const float* chol = nullptr;
};

struct StepCfg {
  int n_steps;
  int step0;
  float dt;
  int n_leapfrog;
  int thinning;
};

template <int NV>
__device__ __forceinline__ void load_row(const float* __restrict__ src, int d, int r, int lpc,
                                         float4 (&x)[NV]) {
  const bool vec = (d & 3) == 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    int j0 = (v * lpc + r) * 4;
    if (vec) {
      x[v] = (j0 < d) ? *reinterpret_cast<const float4*>(src + j0) : make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) f4(x[v], e) = (j0 + e < d) ? src[j0 + e] : 0.f;
    }
  }
}

template <int NV>
__device__ __forceinline__ void store_row(float* __restrict__ dst, int d, int r, int lpc,
                                          const float4 (&x)[NV]) {
  const bool vec = (d & 3) == 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    int j0 = (v * lpc + r) * 4;
    if (vec) {
      if (j0 < d) *reinterpret_cast<float4*>(dst + j0) = x[v];
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (j0 + e < d) dst[j0 + e] = f4c(x[v], e);
    }
  }
}

// Standard normals of (chain, step) for this lane's chunks; padding zeroed.
template <int NV>
__device__ __forceinline__ void draw_normals(const ChainIO& io, int d, int r, int lpc, int64_t c,
                                             int s, int t_glob, uint32_t sid, float4 (&xi)[NV]) {
  if (io.noise) {
    load_row<NV>(io.noise + ((int64_t)s * io.C + c) * d, d, r, lpc, xi);
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int chunk = v * lpc + r;
      xi[v] = philox_normal4(io.key0, io.key1, (uint32_t)chunk, (uint32_t)c, (uint32_t)t_glob, sid);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (chunk * 4 + e >= d) f4(xi[v], e) = 0.f;
    }
  }
}

__device__ __forceinline__ float draw_log_u(const ChainIO& io, int64_t c, int s, int t_glob,
                                            uint32_t sid) {
  if (io.log_u) return io.log_u[(int64_t)s * io.C + c];
  return philox_log_uniform(io.key0, io.key1, (uint32_t)c, (uint32_t)t_glob, sid);
}

template <int NV>
__device__ __forceinline__ void write_rows(const ChainIO& io, int d, int r, int lpc, int64_t c,
                                           int t_glob, int thinning, const float4 (&x)[NV], float lp,
                                           bool acc) {
  if (t_glob % thinning != 0) return;
  int64_t row = io.cursor + t_glob / thinning;
  if (io.out_pos) store_row<NV>(io.out_pos + (c * io.buf_len + row) * d, d, r, lpc, x);
  if (r == 0) {
    if (io.out_logp) io.out_logp[c * io.buf_len + row] = lp;
    if (io.out_acc) io.out_acc[c * io.buf_len + row] = acc ? 1.f : 0.f;
  }
}

// Tempered rows (parallel tempering): row c of a tempered launch targets
// log p(x) / T with T = temps[c % n_temps] — TemperedPDF.tempered_log_pdf
// (src/flowMC/resource/logPDF.py:104-106) with a zero log prior. The value and
// the gradient of the target are scaled by beta = 1/T after each evaluation.
__device__ __forceinline__ float row_beta(const ChainIO& io, int64_t c) {
  return io.temps ? 1.f / io.temps[c % io.n_temps] : 1.f;
}

template <int NV>
__device__ __forceinline__ float temper(const ChainIO& io, float beta, float lp, float4 (&g)[NV]) {
  if (io.temps == nullptr) return lp;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) f4(g[v], e) *= beta;
  return beta * lp;
}

// ------------------------------------------------------------------ MALA ---
// One chain group runs cfg.n_steps MALA steps (src/flowMC/resource/kernel/
// MALA.py:26-89). The gradient at the current point is carried from the
// previous step instead of being re-evaluated: MALA.py:59 recomputes it at the
// same point, which for a deterministic target is the same value.
template <int NV, int TK>
__device__ __forceinline__ void mala_group(const TargetDev& t, const ChainIO& io, const StepCfg& cfg, int64_t c,
                           int r, int lpc, int gbase) {
  const int d = t.d;
  float4 x[NV], g[NV];
  load_row<NV>(io.x + c * d, d, r, lpc, x);
  const float beta = row_beta(io, c);
  float lp = temper<NV>(io, beta, target_vg<NV, TK>(t, x, g, r, lpc, gbase), g);
  const float dt = cfg.dt;
  const float h = 0.5f * dt * dt;
  const float inv2dt2 = 1.f / (2.f * dt * dt);
  for (int s = 0; s < cfg.n_steps; ++s) {
    const int tg = cfg.step0 + s;
    float4 xi[NV], prop[NV], g1[NV];
    draw_normals<NV>(io, d, r, lpc, c, s, tg, RNG_MALA_NOISE, xi);
    float s1 = 0.f;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float n = dt * f4c(xi[v], e);
        f4(prop[v], e) = f4c(x[v], e) + h * f4c(g[v], e) + n;
        s1 += n * n;
      }
    float lp1 = temper<NV>(io, beta, target_vg<NV, TK>(t, prop, g1, r, lpc, gbase), g1);
    float s2 = 0.f;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float q = f4c(x[v], e) - f4c(prop[v], e) - h * f4c(g1[v], e);
        s2 += q * q;
      }
    // padding lanes contribute exactly 0 to s1 and s2
    float ds = group_sum(s1 - s2, lpc);
    float ratio = (lp1 - lp) + ds * inv2dt2;
    float lu = draw_log_u(io, c, s, tg, RNG_MALA_UNIFORM);
    bool acc = lu < ratio;
    if (acc) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        x[v] = prop[v];
        g[v] = g1[v];
      }
      lp = lp1;
    }
    write_rows<NV>(io, d, r, lpc, c, tg, cfg.thinning, x, lp, acc);
  }
  store_row<NV>(io.x + c * d, d, r, lpc, x);
  if (r == 0 && io.logp) io.logp[c] = lp;
}

// ------------------------------------------------------------------- GRW ---
// GaussianRandomWalk.kernel (src/flowMC/resource/kernel/Gaussian_random_walk.py:
// 25-61) scanned by TakeSerialSteps: x' = x + step * xi, accept iff
// log u < logp(x') - logp(x). The log-density of the entry state is evaluated
// once (take_steps.py:177) and carried; the target's gradient output is dead
// and removed by the compiler.
template <int NV, int TK>
__device__ __forceinline__ void grw_group(const TargetDev& t, const ChainIO& io, const StepCfg& cfg, int64_t c,
                          int r, int lpc, int gbase) {
  const int d = t.d;
  float4 x[NV], g[NV];
  load_row<NV>(io.x + c * d, d, r, lpc, x);
  const float beta = row_beta(io, c);
  float lp = temper<NV>(io, beta, target_vg<NV, TK>(t, x, g, r, lpc, gbase), g);
  const float dt = cfg.dt;
  for (int s = 0; s < cfg.n_steps; ++s) {
    const int tg = cfg.step0 + s;
    float4 xi[NV], prop[NV];
    draw_normals<NV>(io, d, r, lpc, c, s, tg, RNG_GRW_NOISE, xi);
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) f4(prop[v], e) = f4c(x[v], e) + dt * f4c(xi[v], e);
    const float lp1 = temper<NV>(io, beta, target_vg<NV, TK>(t, prop, g, r, lpc, gbase), g);
    const float lu = draw_log_u(io, c, s, tg, RNG_GRW_UNIFORM);
    const bool acc = lu < lp1 - lp;
    if (acc) {
#pragma unroll
      for (int v = 0; v < NV; ++v) x[v] = prop[v];
      lp = lp1;
    }
    write_rows<NV>(io, d, r, lpc, c, tg, cfg.thinning, x, lp, acc);
  }
  store_row<NV>(io.x + c * d, d, r, lpc, x);
  if (r == 0 && io.logp) io.logp[c] = lp;
}

// ------------------------------------------------------------------- HMC ---
// HMC.kernel (src/flowMC/resource/kernel/HMC.py:98-151) with the leapfrog
// coefficient table of HMC.py:47-50: row 0 = (0, 1/2), rows 1..L = (1, 1),
// row L+1 = (1, 1/2): L+1 drifts x += eps*c0*w*p and L+2 kicks
// p += eps*c1*grad logp(x). H0 - H1 is accumulated in FP64.
//
// Metric kinds MK (the condition matrix M):
//   MK_UNIT  identity (io.metric == nullptr): padding elements of p, x and
//            the gradient are 0 throughout, so this form needs no
//            per-element index test and no metric multiply;
//   MK_DIAG  diagonal: p = xi / sqrt(m), kinetic 1/2 sum m p^2, drift m p;
//   MK_DENSE dense: p = xi @ chol(inv(M))^T (HMC.py:133-136), and the
//            reference's broadcast kinetic 1/2 sum(p^2 * M) =
//            1/2 sum_j w_j p_j^2 with w = colsum(M) (HMC.py:118-121), whose
//            p-gradient w p is the drift; io.chol holds the factor, io.metric w.
enum { MK_UNIT = 0, MK_DIAG = 1, MK_DENSE = 2 };

// p = L xi for the chain held by this lane group (L lower-triangular [d, d]
// row-major): every lane needs all of its chain's xi, gathered chunk by
// chunk from the group's lanes with shuffles; padding elements of p are 0.
template <int NV>
__device__ __forceinline__ void dense_momentum(const float* __restrict__ Lc, int d, int r, int lpc, int gbase,
                                               const float4 (&xi)[NV], float4 (&p)[NV]) {
#pragma unroll
  for (int v = 0; v < NV; ++v) p[v] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int src = 0; src < lpc; ++src) {
#pragma unroll
    for (int vs = 0; vs < NV; ++vs) {
      const int k0 = (vs * lpc + src) * 4;  // the same for every lane of the wave
      float xs[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) xs[e] = __shfl(f4c(xi[vs], e), gbase + src, 64);
      if (k0 >= d) continue;
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int j = dim_of(v, e, r, lpc);
          if (j < d && k0 <= j) {
            const float* row = Lc + (int64_t)j * d + k0;
            float acc = f4c(p[v], e);
#pragma unroll
            for (int ee = 0; ee < 4; ++ee)
              if (k0 + ee <= j) acc += row[ee] * xs[ee];
            f4(p[v], e) = acc;
          }
        }
    }
  }
}

template <int NV, int MK, int TK>
__device__ __forceinline__ void hmc_group(const TargetDev& t, const ChainIO& io, const StepCfg& cfg, int64_t c,
                          int r, int lpc, int gbase) {
  constexpr bool UNIT_M = MK == MK_UNIT;
  const int d = t.d;
  float4 x[NV], g[NV], m[NV];
  load_row<NV>(io.x + c * d, d, r, lpc, x);
  if (!UNIT_M) {
    load_row<NV>(io.metric, d, r, lpc, m);
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) m[v] = make_float4(1.f, 1.f, 1.f, 1.f);
  }
  const float beta = row_beta(io, c);
  float lp = temper<NV>(io, beta, target_vg<NV, TK>(t, x, g, r, lpc, gbase), g);
  const float eps = cfg.dt;
  const int L = cfg.n_leapfrog;
  for (int s = 0; s < cfg.n_steps; ++s) {
    const int tg = cfg.step0 + s;
    float4 p[NV], xl[NV], gl[NV];
    draw_normals<NV>(io, d, r, lpc, c, s, tg, RNG_HMC_MOMENTUM, p);
    if (MK == MK_DENSE) {
      float4 xi[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) xi[v] = p[v];
      dense_momentum<NV>(io.chol, d, r, lpc, gbase, xi, p);
    }
    double k0 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (UNIT_M) {
          // sum of squares; the 1/2 is applied after the group sum (a power
          // of two: the same bits as summing 0.5 p^2)
          const double pj = (double)f4c(p[v], e);
          k0 = fma(pj, pj, k0);
        } else {
          int j = dim_of(v, e, r, lpc);
          float mj = f4c(m[v], e);
          float pj = (j < d) ? (MK == MK_DENSE ? f4c(p[v], e) : f4c(p[v], e) / sqrtf(mj)) : 0.f;
          f4(p[v], e) = pj;
          k0 += 0.5 * (double)mj * (double)pj * (double)pj;
        }
      }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      xl[v] = x[v];
      gl[v] = g[v];
    }
    k0 = group_sum(k0, lpc);
    if (UNIT_M) k0 *= 0.5;
    // row 0: drift coefficient 0, kick 1/2
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) f4(p[v], e) += eps * 0.5f * f4c(gl[v], e);
    float lpl = lp;
    for (int i = 1; i <= L + 1; ++i) {
      const float kick = (i <= L) ? 1.f : 0.5f;
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) f4(xl[v], e) += eps * (UNIT_M ? f4c(p[v], e) : f4c(m[v], e) * f4c(p[v], e));
      lpl = temper<NV>(io, beta, target_vg<NV, TK>(t, xl, gl, r, lpc, gbase), gl);
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) f4(p[v], e) += eps * kick * f4c(gl[v], e);
    }
    double k1 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (UNIT_M) {
          const double pj = (double)f4c(p[v], e);
          k1 = fma(pj, pj, k1);
        } else {
          int j = dim_of(v, e, r, lpc);
          if (j < d) k1 += 0.5 * (double)f4c(m[v], e) * (double)f4c(p[v], e) * (double)f4c(p[v], e);
        }
      }
    k1 = group_sum(k1, lpc);
    if (UNIT_M) k1 *= 0.5;
    double h0 = -(double)lp + k0;
    double h1 = -(double)lpl + k1;
    double log_acc = h0 - h1;
    float lu = draw_log_u(io, c, s, tg, RNG_HMC_UNIFORM);
    bool acc = (double)lu < log_acc;
    if (acc) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        x[v] = xl[v];
        g[v] = gl[v];
      }
      lp = lpl;
    }
    write_rows<NV>(io, d, r, lpc, c, tg, cfg.thinning, x, lp, acc);
  }
  store_row<NV>(io.x + c * d, d, r, lpc, x);
  if (r == 0 && io.logp) io.logp[c] = lp;
}

// Copies the target parameter block to LDS when it fits (n_lds > 0).
__device__ __forceinline__ TargetDev stage_target(const fmc_target& tg, float* smem, int n_lds) {
  TargetDev t{tg.kind, tg.n_dim, tg.n_modes, tg.params};
  if (n_lds > 0) {
    for (int i = threadIdx.x; i < n_lds; i += blockDim.x) smem[i] = tg.params[i];
    __syncthreads();
    t.p = smem;
  }
  return t;
}

}  // namespace fmc
