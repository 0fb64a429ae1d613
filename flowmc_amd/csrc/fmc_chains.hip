// Chain kernels for gfx950: fused K-step MALA / HMC over built-in targets,
// per-step MALA/HMC pieces for externally differentiated targets, the
// normalizing-flow independence MH scan, and the RNG streams they use.
//
// Work mapping: a chain is owned by `lpc` lanes of one 64-wide wavefront
// (lpc a power of two <= 64, groups aligned inside the wave); each lane keeps
// NV float4 chunks of position / gradient in registers for the whole launch,
// so the K fused steps touch HBM only for the thinned buffer rows, the noise
// (injected-noise mode) and the state at entry/exit.
#include "fmc_chain_kernel.h"

namespace fmc {

// ------------------------------------------------------- target eval ------
template <int NV>
__global__ void __launch_bounds__(1024) target_eval_kernel(fmc_target tg, const float* __restrict__ x,
                                                           int64_t n, float* logp, float* grad,
                                                           int lpc, int n_lds) {
  extern __shared__ float smem[];
  TargetDev t = stage_target(tg, smem, n_lds);
  const int lane = threadIdx.x & 63;
  const int r = lane & (lpc - 1);
  const int gbase = lane - r;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x / lpc) + threadIdx.x / lpc;
  if (c >= n) return;
  const int d = t.d;
  float4 xv[NV], g[NV];
  load_row<NV>(x + c * d, d, r, lpc, xv);
  float lp = target_vg<NV>(t, xv, g, r, lpc, gbase);
  if (grad) store_row<NV>(grad + c * d, d, r, lpc, g);
  if (r == 0 && logp) logp[c] = lp;
}

// ------------------------------------------- per-step pieces (any target) --
// One 64-lane wave per chain; lanes stride over dimensions.
__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float noise_at(const float* noise, uint32_t k0, uint32_t k1, int64_t c,
                                          int64_t C, int d, int j, int step, uint32_t sid) {
  if (noise) return noise[c * d + j];
  // same (chunk, chain, step, stream) counter as the fused kernels
  float4 z = philox_normal4(k0, k1, (uint32_t)(j >> 2), (uint32_t)c, (uint32_t)step, sid);
  return f4c(z, j & 3);
}

__global__ void mala_propose_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                    float* __restrict__ prop, int64_t C, int d, float dt,
                                    uint32_t k0, uint32_t k1, int step, const float* noise) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C * d) return;
  const int64_t c = i / d;
  const int j = (int)(i - c * d);
  float xi = noise_at(noise, k0, k1, c, C, d, j, step, RNG_MALA_NOISE);
  prop[i] = x[i] + 0.5f * dt * dt * g[i] + dt * xi;
}

__global__ void mala_accept_kernel(float* x, float* logp, float* g, const float* __restrict__ prop,
                                   const float* __restrict__ lp1, const float* __restrict__ g1,
                                   int64_t C, int d, float dt, uint32_t k0, uint32_t k1, int step,
                                   const float* log_u, float* out_pos, float* out_logp,
                                   float* out_acc, int64_t buf_len, int64_t cursor, int thinning) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= C) return;
  const float h = 0.5f * dt * dt;
  float s1 = 0.f, s2 = 0.f;
  for (int j = lane; j < d; j += 64) {
    int64_t k = c * d + j;
    float a = prop[k] - x[k] - h * g[k];
    float b = x[k] - prop[k] - h * g1[k];
    s1 += a * a;
    s2 += b * b;
  }
  float ds = wave_sum(s1 - s2);
  float ratio = (lp1[c] - logp[c]) + ds / (2.f * dt * dt);
  float lu = log_u ? log_u[c] : philox_log_uniform(k0, k1, (uint32_t)c, (uint32_t)step, RNG_MALA_UNIFORM);
  bool acc = lu < ratio;
  const bool wr = (step % thinning) == 0;
  const int64_t row = cursor + step / thinning;
  for (int j = lane; j < d; j += 64) {
    int64_t k = c * d + j;
    float xv = acc ? prop[k] : x[k];
    if (acc) {
      x[k] = xv;
      g[k] = g1[k];
    }
    if (wr && out_pos) out_pos[(c * buf_len + row) * d + j] = xv;
  }
  if (lane == 0) {
    float lpv = acc ? lp1[c] : logp[c];
    logp[c] = lpv;
    if (wr && out_logp) out_logp[c * buf_len + row] = lpv;
    if (wr && out_acc) out_acc[c * buf_len + row] = acc ? 1.f : 0.f;
  }
}

// Gaussian random walk pieces for targets evaluated outside the library.
__global__ void grw_propose_kernel(const float* __restrict__ x, float* __restrict__ prop, int64_t C, int d,
                                   float dt, uint32_t k0, uint32_t k1, int step, const float* noise) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C * d) return;
  const int64_t c = i / d;
  const int j = (int)(i - c * d);
  prop[i] = x[i] + dt * noise_at(noise, k0, k1, c, C, d, j, step, RNG_GRW_NOISE);
}

__global__ void grw_accept_kernel(float* x, float* logp, const float* __restrict__ prop,
                                  const float* __restrict__ lp1, int64_t C, int d, uint32_t k0, uint32_t k1,
                                  int step, const float* log_u, float* out_pos, float* out_logp,
                                  float* out_acc, int64_t buf_len, int64_t cursor, int thinning) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= C) return;
  const float lu = log_u ? log_u[c] : philox_log_uniform(k0, k1, (uint32_t)c, (uint32_t)step, RNG_GRW_UNIFORM);
  const bool acc = lu < lp1[c] - logp[c];
  const bool wr = (step % thinning) == 0;
  const int64_t row = cursor + step / thinning;
  for (int j = lane; j < d; j += 64) {
    const int64_t k = c * d + j;
    const float xv = acc ? prop[k] : x[k];
    if (acc) x[k] = xv;
    if (wr && out_pos) out_pos[(c * buf_len + row) * d + j] = xv;
  }
  if (lane == 0) {
    const float lpv = acc ? lp1[c] : logp[c];
    logp[c] = lpv;
    if (wr && out_logp) out_logp[c * buf_len + row] = lpv;
    if (wr && out_acc) out_acc[c * buf_len + row] = acc ? 1.f : 0.f;
  }
}

// One wavefront per chain. Dense condition matrix (chol != NULL): the draw xi
// is staged in LDS (d floats per wavefront) and p = L xi with L = chol(inv(M))
// lower-triangular; metric then holds the kinetic weights colsum(M).
__global__ void hmc_momentum_kernel(float* p, double* kinetic, int64_t C, int d, const float* metric,
                                    const float* chol, uint32_t k0, uint32_t k1, int step, const float* noise) {
  extern __shared__ float xi_lds[];
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= C) return;
  double k = 0.0;
  if (chol) {
    float* xs = xi_lds + (threadIdx.x >> 6) * d;
    for (int j = lane; j < d; j += 64) xs[j] = noise_at(noise, k0, k1, c, C, d, j, step, RNG_HMC_MOMENTUM);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (int j = lane; j < d; j += 64) {
      const float* row = chol + (int64_t)j * d;
      float pj = 0.f;
      for (int kk = 0; kk <= j; ++kk) pj += row[kk] * xs[kk];
      const float wj = metric[j];
      p[c * d + j] = pj;
      k += 0.5 * (double)wj * (double)pj * (double)pj;
    }
  } else {
    for (int j = lane; j < d; j += 64) {
      float mj = metric ? metric[j] : 1.f;
      float pj = noise_at(noise, k0, k1, c, C, d, j, step, RNG_HMC_MOMENTUM) / sqrtf(mj);
      p[c * d + j] = pj;
      k += 0.5 * (double)mj * (double)pj * (double)pj;
    }
  }
  k = wave_sum(k);
  if (lane == 0) kinetic[c] = k;
}

__global__ void hmc_drift_kernel(float* x, const float* __restrict__ p, const float* metric,
                                 int64_t n, int d, float coef) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float mj = metric ? metric[i % d] : 1.f;
  x[i] += coef * (mj * p[i]);
}

__global__ void hmc_kick_kernel(float* p, const float* __restrict__ g, int64_t n, float coef) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] += coef * g[i];
}

__global__ void hmc_accept_kernel(float* x, float* logp, const float* __restrict__ xn,
                                  const float* __restrict__ lpn, const float* __restrict__ pn,
                                  const double* __restrict__ k0, const float* metric, int64_t C,
                                  int d, uint32_t key0, uint32_t key1, int step, const float* log_u,
                                  float* out_pos, float* out_logp, float* out_acc, int64_t buf_len,
                                  int64_t cursor, int thinning) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= C) return;
  double k1 = 0.0;
  for (int j = lane; j < d; j += 64) {
    float mj = metric ? metric[j] : 1.f;
    float pj = pn[c * d + j];
    k1 += 0.5 * (double)mj * (double)pj * (double)pj;
  }
  k1 = wave_sum(k1);
  double log_acc = (-(double)logp[c] + k0[c]) - (-(double)lpn[c] + k1);
  float lu = log_u ? log_u[c] : philox_log_uniform(key0, key1, (uint32_t)c, (uint32_t)step, RNG_HMC_UNIFORM);
  bool acc = (double)lu < log_acc;
  const bool wr = (step % thinning) == 0;
  const int64_t row = cursor + step / thinning;
  for (int j = lane; j < d; j += 64) {
    int64_t k = c * d + j;
    float xv = acc ? xn[k] : x[k];
    x[k] = xv;
    if (wr && out_pos) out_pos[(c * buf_len + row) * d + j] = xv;
  }
  if (lane == 0) {
    float lpv = acc ? lpn[c] : logp[c];
    logp[c] = lpv;
    if (wr && out_logp) out_logp[c * buf_len + row] = lpv;
    if (wr && out_acc) out_acc[c * buf_len + row] = acc ? 1.f : 0.f;
  }
}

// -------------------------------------------------------------- RNG -------
// ---------------------------------------------------- finite-row mask ---
// out[r] = 1 iff every entry of row r is finite (jnp.isfinite(data).all(-1)
// in TrainModel, src/flowMC/strategy/train_model.py:66-80). A row is read by
// g lanes of one wavefront (g a power of two), float4-wide when vec is set;
// one ballot per wavefront combines the lanes of each row. The grid strides
// over passes of 256 / g rows, FR_UNROLL passes per iteration, so every lane
// has that many loads in flight (the buffer is the whole sample history: a
// streaming read of gigabytes). On the float4 path the first lane of each
// row reads the row's first four entries ahead of the others: a row that
// fails there (the unfilled rows of a sample buffer hold -inf) is decided
// without reading the rest of it.
constexpr int FR_UNROLL = 4;
__device__ __forceinline__ bool finite4(const float4 v) {
  return isfinite(v.x) & isfinite(v.y) & isfinite(v.z) & isfinite(v.w);
}
__global__ void __launch_bounds__(256) finite_rows_kernel(const float* __restrict__ x, int64_t n_rows, int d,
                                                          int g, int vec, int32_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int sub = lane & (g - 1);
  const int rpb = 256 / g;  // rows per workgroup pass
  const int64_t stride = (int64_t)gridDim.x * rpb;
  const unsigned long long grp = (g == 64) ? ~0ull : (((1ull << g) - 1ull) << (lane & ~(g - 1)));
  const int lead = lane & ~(g - 1);  // the row's first lane
  // base is uniform over the workgroup, so every wavefront reaches each ballot
  for (int64_t base = (int64_t)blockIdx.x * rpb; base < n_rows; base += FR_UNROLL * stride) {
    bool ok[FR_UNROLL];
    if (vec) {
      bool head[FR_UNROLL];
#pragma unroll
      for (int u = 0; u < FR_UNROLL; ++u) {
        const int64_t row = base + u * stride + threadIdx.x / g;
        bool h = true;
        if (sub == 0 && row < n_rows) h = finite4(*reinterpret_cast<const float4*>(x + row * d));
        head[u] = h;
      }
#pragma unroll
      for (int u = 0; u < FR_UNROLL; ++u) {
        const int64_t row = base + u * stride + threadIdx.x / g;
        const bool dead = (__ballot(!head[u]) >> lead) & 1ull;
        bool o = !dead;
        if (!dead && row < n_rows) {
          const float* p = x + row * d;
          for (int j = sub * 4; j < d; j += g * 4) o &= finite4(*reinterpret_cast<const float4*>(p + j));
        }
        ok[u] = o;
      }
    } else {
#pragma unroll
      for (int u = 0; u < FR_UNROLL; ++u) {
        const int64_t row = base + u * stride + threadIdx.x / g;
        bool o = true;
        if (row < n_rows) {
          const float* p = x + row * d;
          for (int j = sub; j < d; j += g) o &= isfinite(p[j]);
        }
        ok[u] = o;
      }
    }
#pragma unroll
    for (int u = 0; u < FR_UNROLL; ++u) {
      const int64_t row = base + u * stride + threadIdx.x / g;
      const unsigned long long bad = __ballot(!ok[u]);
      if (sub == 0 && row < n_rows) out[row] = (bad & grp) ? 0 : 1;
    }
  }
}

// --------------------------------------------------- tempering exchange ---
// ParallelTempering._exchange (src/flowMC/strategy/parallel_tempering.py:
// 291-372), PT_CPB chains per block. Per chain, the pairs (i, i+1),
// i = 0 .. T-2, are visited in order and the swap is accepted iff
//   log u < (1/T[i+1] - 1/T[i]) * (L_i - L_{i+1})
// with the log-likelihoods of the walkers currently in slots i and i+1, so a
// walker can climb several slots in one sweep. Once pair i is decided slot i
// is final, so the scan carries only the walker that may still move up (its
// source slot and log-likelihood). Phase 1 (one thread per chain) writes the
// slot -> source permutation to LDS and the log-likelihoods and decisions to
// memory; phase 2 gathers the rows, one wave per row.
constexpr int PT_CPB = 64;
constexpr int PT_TMAX = 64;

__global__ void __launch_bounds__(256) pt_exchange_kernel(const float* __restrict__ x_in, float* __restrict__ x_out,
                                                          float* loglik, const float* __restrict__ temps,
                                                          int64_t C, int T, int d, uint32_t k0, uint32_t k1,
                                                          const float* __restrict__ log_u,
                                                          float* __restrict__ acc) {
  __shared__ int src[PT_CPB * PT_TMAX];
  const int64_t c0 = (int64_t)blockIdx.x * PT_CPB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nc = (C - c0) < PT_CPB ? (int)(C - c0) : PT_CPB;
  if (tid < nc) {
    const int64_t c = c0 + tid;
    float* L = loglik + c * T;
    int* sl = src + tid * PT_TMAX;
    int cur = 0;
    float lc = L[0];
    for (int i = 0; i + 1 < T; ++i) {
      const float ln = L[i + 1];
      const float ratio = (1.f / temps[i + 1] - 1.f / temps[i]) * (lc - ln);
      const float lu = log_u ? log_u[c * (T - 1) + i]
                             : philox_log_uniform(k0, k1, (uint32_t)c, (uint32_t)i, RNG_PT_SWAP);
      const bool a = lu < ratio;
      acc[c * (T - 1) + i] = a ? 1.f : 0.f;
      if (a) {  // slot i takes walker i+1, the carried walker moves up to slot i+1
        sl[i] = i + 1;
        L[i] = ln;
      } else {  // slot i keeps the carried walker, walker i+1 is carried on
        sl[i] = cur;
        L[i] = lc;
        cur = i + 1;
        lc = ln;
      }
    }
    sl[T - 1] = cur;
    L[T - 1] = lc;
  }
  __syncthreads();
  for (int row = wave; row < nc * T; row += 4) {
    const int cl = row / T, t = row - cl * T;
    const int64_t base = (c0 + cl) * T;
    const float* s = x_in + (base + src[cl * PT_TMAX + t]) * d;
    float* o = x_out + (base + t) * d;
    for (int j = lane; j < d; j += 64) o[j] = s[j];
  }
}

// ParallelTempering._adapt_temperature (parallel_tempering.py:400-436), in
// place, one block: a_p = mean over chains of the swap decisions of pair p;
// damping_p = (100 / C) (a_p - a_{p+1});
// T'[i] = T'[i-1] + (T[i] - T[i-1]) exp(damping_{i-1}) for i = 1 .. T-2;
// the first and last temperatures are unchanged.
__global__ void __launch_bounds__(256) pt_adapt_kernel(float* temps, const float* __restrict__ acc, int64_t C,
                                                       int T) {
  __shared__ float rate[PT_TMAX];
  __shared__ float part[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int p = 0; p + 1 < T; ++p) {
    float s = 0.f;  // sums of 0/1 values: exact below 2^24 per partial
    for (int64_t c = tid; c < C; c += 256) s += acc[c * (T - 1) + p];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) part[wave] = s;
    __syncthreads();
    if (tid == 0) rate[p] = ((part[0] + part[1]) + (part[2] + part[3])) / (float)C;
    __syncthreads();
  }
  if (tid == 0) {
    const float k = 100.f / (float)C;
    float prev_new = temps[0], prev_old = temps[0];
    for (int i = 1; i + 1 < T; ++i) {
      const float old = temps[i];
      const float nw = prev_new + (old - prev_old) * expf(k * (rate[i - 1] - rate[i]));
      temps[i] = nw;
      prev_new = nw;
      prev_old = old;
    }
  }
}

__global__ void philox_blocks_kernel(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                                     uint32_t c3, int64_t n, uint32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32x4 b = philox4x32_10(u32x4{c0 + (uint32_t)i, c1, c2, c3}, k0, k1);
  out[4 * i + 0] = b.x;
  out[4 * i + 1] = b.y;
  out[4 * i + 2] = b.z;
  out[4 * i + 3] = b.w;
}

__global__ void rng_normal_kernel(uint32_t k0, uint32_t k1, int64_t C, int d, int n_steps, int step0,
                                  uint32_t sid, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_steps * C * d) return;
  const int64_t row = i / d;
  const int j = (int)(i - row * d);
  const int s = (int)(row / C);
  const int64_t c = row - (int64_t)s * C;
  float4 z = philox_normal4(k0, k1, (uint32_t)(j >> 2), (uint32_t)c, (uint32_t)(step0 + s), sid);
  out[i] = f4c(z, j & 3);
}

__global__ void rng_log_uniform_kernel(uint32_t k0, uint32_t k1, int64_t C, int n_steps, int step0,
                                       uint32_t sid, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_steps * C) return;
  const int s = (int)(i / C);
  const int64_t c = i - (int64_t)s * C;
  out[i] = philox_log_uniform(k0, k1, (uint32_t)c, (uint32_t)(step0 + s), sid);
}

}  // namespace fmc

using namespace fmc;

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

extern "C" int fmc_target_check(const fmc_target* target) { return check_target(target); }

extern "C" int fmc_target_eval(void* stream, const fmc_target* target, const float* x, int64_t n_points,
                               float* logp, float* grad, const fmc_launch_cfg* cfg) {
  int rc = check_target(target);
  if (rc) return rc;
  FMC_ARG(x != nullptr && n_points >= 0, "x is NULL or n_points < 0");
  if (n_points == 0) return FMC_OK;
  Layout L = choose_layout(target->n_dim, target->kind, cfg);
  if ((rc = check_layout(L, target->n_dim))) return rc;
  const int per_block = L.block / L.lpc;
  const int64_t grid = (n_points + per_block - 1) / per_block;
  const int n_lds = lds_floats(target);
  const size_t shm = (size_t)n_lds * sizeof(float);
  hipStream_t st = (hipStream_t)stream;
  dim3 g((unsigned)grid), b((unsigned)L.block);
  switch (L.nv) {
    case 1: hipLaunchKernelGGL(target_eval_kernel<1>, g, b, shm, st, *target, x, n_points, logp, grad, L.lpc, n_lds); break;
    case 2: hipLaunchKernelGGL(target_eval_kernel<2>, g, b, shm, st, *target, x, n_points, logp, grad, L.lpc, n_lds); break;
    case 4: hipLaunchKernelGGL(target_eval_kernel<4>, g, b, shm, st, *target, x, n_points, logp, grad, L.lpc, n_lds); break;
    default: hipLaunchKernelGGL(target_eval_kernel<8>, g, b, shm, st, *target, x, n_points, logp, grad, L.lpc, n_lds); break;
  }
  return check_launch("target eval kernel");
}

extern "C" int fmc_mala_run(void* stream, const fmc_target* target, float* x, float* logp, int64_t n_chains,
                            int32_t n_steps, int32_t step0, float step_size, uint32_t key0, uint32_t key1,
                            const float* noise, const float* log_u, float* out_pos, float* out_logp,
                            float* out_acc, int64_t buf_len, int64_t cursor, int32_t thinning,
                            const fmc_launch_cfg* cfg) {
  int rc = check_target(target);
  if (rc) return rc;
  FMC_ARG(x != nullptr, "x is NULL");
  FMC_ARG(n_chains >= 0, "n_chains < 0");
  FMC_ARG(step_size > 0.f, "step_size must be > 0");
  FMC_ARG((noise == nullptr) == (log_u == nullptr), "noise and log_u must be both given or both NULL");
  if ((rc = check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, step0, n_steps))) return rc;
  if (n_chains == 0) return FMC_OK;
  Layout L = choose_layout(target->n_dim, target->kind, cfg);
  if ((rc = check_layout(L, target->n_dim))) return rc;
  ChainIO io{x, logp, nullptr, noise, log_u, out_pos, out_logp, out_acc, n_chains, buf_len, cursor, key0, key1};
  StepCfg sc{n_steps, step0, step_size, 0, thinning};
  return launch_chain<KK_MALA>((hipStream_t)stream, target, io, sc, L);
}

extern "C" int fmc_grw_run(void* stream, const fmc_target* target, float* x, float* logp, int64_t n_chains,
                           int32_t n_steps, int32_t step0, float step_size, uint32_t key0, uint32_t key1,
                           const float* noise, const float* log_u, float* out_pos, float* out_logp,
                           float* out_acc, int64_t buf_len, int64_t cursor, int32_t thinning,
                           const fmc_launch_cfg* cfg) {
  int rc = check_target(target);
  if (rc) return rc;
  FMC_ARG(x != nullptr, "x is NULL");
  FMC_ARG(n_chains >= 0, "n_chains < 0");
  FMC_ARG(step_size > 0.f, "step_size must be > 0");
  FMC_ARG((noise == nullptr) == (log_u == nullptr), "noise and log_u must be both given or both NULL");
  if ((rc = check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, step0, n_steps))) return rc;
  if (n_chains == 0) return FMC_OK;
  Layout L = choose_layout(target->n_dim, target->kind, cfg);
  if ((rc = check_layout(L, target->n_dim))) return rc;
  ChainIO io{x, logp, nullptr, noise, log_u, out_pos, out_logp, out_acc, n_chains, buf_len, cursor, key0, key1};
  StepCfg sc{n_steps, step0, step_size, 0, thinning};
  return launch_chain<KK_GRW>((hipStream_t)stream, target, io, sc, L);
}

extern "C" int fmc_tempered_run(void* stream, int32_t kernel, const fmc_target* target, float* x, float* logp,
                                int64_t n_chains, int32_t n_temps, const float* temperatures, int32_t n_steps,
                                int32_t step0, float step_size, int32_t n_leapfrog, const float* metric,
                                const float* momentum_chol, uint32_t key0, uint32_t key1, const float* noise,
                                const float* log_u, float* out_acc, int64_t acc_len, const fmc_launch_cfg* cfg) {
  int rc = check_target(target);
  if (rc) return rc;
  FMC_ARG(kernel == FMC_KERNEL_MALA || kernel == FMC_KERNEL_HMC || kernel == FMC_KERNEL_GRW,
          "unknown kernel %d", kernel);
  FMC_ARG(x != nullptr && temperatures != nullptr, "x or temperatures is NULL");
  FMC_ARG(n_chains >= 0 && n_temps >= 1, "n_chains must be >= 0 and n_temps >= 1");
  FMC_ARG(kernel == FMC_KERNEL_HMC || step_size > 0.f, "step_size must be > 0");
  FMC_ARG(n_leapfrog >= 0, "n_leapfrog < 0");
  FMC_ARG((noise == nullptr) == (log_u == nullptr), "noise and log_u must be both given or both NULL");
  if ((rc = check_buffers(nullptr, nullptr, out_acc, acc_len, 0, 1, step0, n_steps))) return rc;
  const int64_t rows = n_chains * n_temps;
  if (rows == 0) return FMC_OK;
  Layout L = choose_layout(target->n_dim, target->kind, cfg);
  if ((rc = check_layout(L, target->n_dim))) return rc;
  FMC_ARG(momentum_chol == nullptr || (kernel == FMC_KERNEL_HMC && metric != nullptr),
          "a momentum Cholesky factor needs the HMC kernel and the kinetic weights in metric");
  ChainIO io{x, logp, kernel == FMC_KERNEL_HMC ? metric : nullptr, noise, log_u, nullptr, nullptr, out_acc,
             rows, acc_len, 0, key0, key1, temperatures, n_temps};
  io.chol = kernel == FMC_KERNEL_HMC ? momentum_chol : nullptr;
  StepCfg sc{n_steps, step0, step_size, n_leapfrog, 1};
  hipStream_t st = (hipStream_t)stream;
  if (kernel == FMC_KERNEL_MALA) return launch_chain<KK_MALA>(st, target, io, sc, L);
  if (kernel == FMC_KERNEL_GRW) return launch_chain<KK_GRW>(st, target, io, sc, L);
  return launch_chain<KK_HMC>(st, target, io, sc, L);
}

extern "C" int fmc_finite_rows(void* stream, const float* data, int64_t n_rows, int32_t n_dim, int32_t* out) {
  FMC_ARG(data != nullptr && out != nullptr, "NULL pointer");
  FMC_ARG(n_rows >= 0 && n_dim >= 1, "bad sizes");
  if (n_rows == 0) return FMC_OK;
  const int vec = ((n_dim & 3) == 0) && (((uintptr_t)data & 15) == 0);
  const int units = vec ? n_dim / 4 : n_dim;
  int g = 1;
  while (g < units && g < 64) g *= 2;
  // eight workgroups per CU on 256 CUs hold the whole grid at once
  int64_t grid = (n_rows * g + 255) / 256;
  if (grid > 2048) grid = 2048;
  hipLaunchKernelGGL(finite_rows_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, data, n_rows,
                     n_dim, g, vec, out);
  return check_launch("finite rows");
}

extern "C" int fmc_pt_exchange(void* stream, const float* x_in, float* x_out, float* loglik,
                               const float* temperatures, int64_t n_chains, int32_t n_temps, int32_t n_dim,
                               uint32_t key0, uint32_t key1, const float* log_u, float* accepts) {
  FMC_ARG(x_in && x_out && loglik && temperatures, "NULL pointer");
  FMC_ARG(x_in != x_out, "x_out must not alias x_in");
  FMC_ARG(n_chains >= 0 && n_dim >= 1, "bad sizes");
  FMC_ARG(n_temps >= 1 && n_temps <= PT_TMAX, "n_temps must be in [1, %d] (got %d)", PT_TMAX, n_temps);
  FMC_ARG(n_temps == 1 || accepts != nullptr, "accepts is NULL");
  if (n_chains == 0) return FMC_OK;
  const int64_t grid = (n_chains + PT_CPB - 1) / PT_CPB;
  FMC_ARG(grid <= 0x7fffffffLL, "too many chains for one launch");
  hipLaunchKernelGGL(pt_exchange_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, x_in, x_out,
                     loglik, temperatures, n_chains, n_temps, n_dim, key0, key1, log_u, accepts);
  return check_launch("tempering exchange");
}

extern "C" int fmc_pt_adapt(void* stream, float* temperatures, const float* accepts, int64_t n_chains,
                            int32_t n_temps) {
  FMC_ARG(temperatures != nullptr, "temperatures is NULL");
  FMC_ARG(n_temps >= 1 && n_temps <= PT_TMAX, "n_temps must be in [1, %d] (got %d)", PT_TMAX, n_temps);
  FMC_ARG(n_chains >= 1, "n_chains must be >= 1");
  if (n_temps < 3) return FMC_OK;  // no interior temperature to move
  FMC_ARG(accepts != nullptr, "accepts is NULL");
  hipLaunchKernelGGL(pt_adapt_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, temperatures, accepts,
                     n_chains, n_temps);
  return check_launch("temperature adaptation");
}

extern "C" int fmc_grw_propose(void* stream, const float* x, float* prop, int64_t n_chains, int32_t n_dim,
                               float step_size, uint32_t key0, uint32_t key1, int32_t step, const float* noise) {
  FMC_ARG(x && prop, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1 && step >= 0, "bad sizes");
  int64_t n = n_chains * n_dim;
  if (n == 0) return FMC_OK;
  hipLaunchKernelGGL(grw_propose_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, x, prop,
                     n_chains, n_dim, step_size, key0, key1, step, noise);
  return check_launch("grw propose");
}

extern "C" int fmc_grw_accept(void* stream, float* x, float* logp, const float* prop, const float* prop_logp,
                              int64_t n_chains, int32_t n_dim, uint32_t key0, uint32_t key1, int32_t step,
                              const float* log_u, float* out_pos, float* out_logp, float* out_acc, int64_t buf_len,
                              int64_t cursor, int32_t thinning) {
  FMC_ARG(x && logp && prop && prop_logp, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1, "bad sizes");
  int rc = check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, step, 1);
  if (rc) return rc;
  if (n_chains == 0) return FMC_OK;
  hipLaunchKernelGGL(grw_accept_kernel, dim3(grid1(n_chains, 4)), dim3(256), 0, (hipStream_t)stream, x, logp,
                     prop, prop_logp, n_chains, n_dim, key0, key1, step, log_u, out_pos, out_logp, out_acc,
                     buf_len, cursor, thinning);
  return check_launch("grw accept");
}

extern "C" int fmc_hmc_run(void* stream, const fmc_target* target, float* x, float* logp, int64_t n_chains,
                           int32_t n_steps, int32_t step0, float step_size, int32_t n_leapfrog,
                           const float* metric, const float* momentum_chol, uint32_t key0, uint32_t key1,
                           const float* noise,
                           const float* log_u, float* out_pos, float* out_logp, float* out_acc,
                           int64_t buf_len, int64_t cursor, int32_t thinning, const fmc_launch_cfg* cfg) {
  int rc = check_target(target);
  if (rc) return rc;
  FMC_ARG(x != nullptr, "x is NULL");
  FMC_ARG(n_chains >= 0, "n_chains < 0");
  FMC_ARG(n_leapfrog >= 0, "n_leapfrog < 0");
  FMC_ARG((noise == nullptr) == (log_u == nullptr), "noise and log_u must be both given or both NULL");
  if ((rc = check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, step0, n_steps))) return rc;
  if (n_chains == 0) return FMC_OK;
  Layout L = choose_layout(target->n_dim, target->kind, cfg);
  if ((rc = check_layout(L, target->n_dim))) return rc;
  FMC_ARG(momentum_chol == nullptr || metric != nullptr,
          "a momentum Cholesky factor needs the kinetic weights in metric");
  ChainIO io{x, logp, metric, noise, log_u, out_pos, out_logp, out_acc, n_chains, buf_len, cursor, key0, key1};
  io.chol = momentum_chol;
  StepCfg sc{n_steps, step0, step_size, n_leapfrog, thinning};
  return launch_chain<KK_HMC>((hipStream_t)stream, target, io, sc, L);
}

extern "C" int fmc_mala_propose(void* stream, const float* x, const float* grad, float* prop, int64_t n_chains,
                                int32_t n_dim, float step_size, uint32_t key0, uint32_t key1, int32_t step,
                                const float* noise) {
  FMC_ARG(x && grad && prop, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1 && step >= 0, "bad sizes");
  int64_t n = n_chains * n_dim;
  if (n == 0) return FMC_OK;
  hipLaunchKernelGGL(mala_propose_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, x, grad,
                     prop, n_chains, n_dim, step_size, key0, key1, step, noise);
  return check_launch("mala propose");
}

extern "C" int fmc_mala_accept(void* stream, float* x, float* logp, float* grad, const float* prop,
                               const float* prop_logp, const float* prop_grad, int64_t n_chains, int32_t n_dim,
                               float step_size, uint32_t key0, uint32_t key1, int32_t step, const float* log_u,
                               float* out_pos, float* out_logp, float* out_acc, int64_t buf_len, int64_t cursor,
                               int32_t thinning) {
  FMC_ARG(x && logp && grad && prop && prop_logp && prop_grad, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1, "bad sizes");
  int rc = check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, step, 1);
  if (rc) return rc;
  if (n_chains == 0) return FMC_OK;
  hipLaunchKernelGGL(mala_accept_kernel, dim3(grid1(n_chains, 4)), dim3(256), 0, (hipStream_t)stream, x, logp,
                     grad, prop, prop_logp, prop_grad, n_chains, n_dim, step_size, key0, key1, step, log_u,
                     out_pos, out_logp, out_acc, buf_len, cursor, thinning);
  return check_launch("mala accept");
}

extern "C" int fmc_hmc_momentum(void* stream, float* p, double* kinetic, int64_t n_chains, int32_t n_dim,
                                const float* metric, const float* momentum_chol, uint32_t key0, uint32_t key1,
                                int32_t step, const float* noise) {
  FMC_ARG(p && kinetic, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1 && step >= 0, "bad sizes");
  FMC_ARG(momentum_chol == nullptr || metric != nullptr, "a momentum Cholesky factor needs the kinetic weights");
  FMC_ARG(momentum_chol == nullptr || n_dim <= 4096, "dense condition matrices need n_dim <= 4096");
  if (n_chains == 0) return FMC_OK;
  const size_t shm = momentum_chol ? (size_t)4 * n_dim * sizeof(float) : 0;
  hipLaunchKernelGGL(hmc_momentum_kernel, dim3(grid1(n_chains, 4)), dim3(256), shm, (hipStream_t)stream, p,
                     kinetic, n_chains, n_dim, metric, momentum_chol, key0, key1, step, noise);
  return check_launch("hmc momentum");
}

extern "C" int fmc_hmc_drift(void* stream, float* x, const float* p, const float* metric, int64_t n_chains,
                             int32_t n_dim, float coef) {
  FMC_ARG(x && p, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1, "bad sizes");
  int64_t n = n_chains * n_dim;
  if (n == 0) return FMC_OK;
  hipLaunchKernelGGL(hmc_drift_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, x, p, metric, n,
                     n_dim, coef);
  return check_launch("hmc drift");
}

extern "C" int fmc_hmc_kick(void* stream, float* p, const float* grad_logp, int64_t n_chains, int32_t n_dim,
                            float coef) {
  FMC_ARG(p && grad_logp, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1, "bad sizes");
  int64_t n = n_chains * n_dim;
  if (n == 0) return FMC_OK;
  hipLaunchKernelGGL(hmc_kick_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, p, grad_logp, n,
                     coef);
  return check_launch("hmc kick");
}

extern "C" int fmc_hmc_accept(void* stream, float* x, float* logp, const float* x_new, const float* logp_new,
                              const float* p_new, const double* kinetic0, const float* metric, int64_t n_chains,
                              int32_t n_dim, uint32_t key0, uint32_t key1, int32_t step, const float* log_u,
                              float* out_pos, float* out_logp, float* out_acc, int64_t buf_len, int64_t cursor,
                              int32_t thinning) {
  FMC_ARG(x && logp && x_new && logp_new && p_new && kinetic0, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_dim >= 1, "bad sizes");
  int rc = check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, step, 1);
  if (rc) return rc;
  if (n_chains == 0) return FMC_OK;
  hipLaunchKernelGGL(hmc_accept_kernel, dim3(grid1(n_chains, 4)), dim3(256), 0, (hipStream_t)stream, x, logp,
                     x_new, logp_new, p_new, kinetic0, metric, n_chains, n_dim, key0, key1, step, log_u, out_pos,
                     out_logp, out_acc, buf_len, cursor, thinning);
  return check_launch("hmc accept");
}

extern "C" int fmc_philox_blocks(void* stream, uint32_t key0, uint32_t key1, uint32_t ctr0, uint32_t ctr1,
                                 uint32_t ctr2, uint32_t ctr3, int64_t n_blocks, uint32_t* out) {
  FMC_ARG(out != nullptr && n_blocks >= 0, "bad output");
  if (n_blocks == 0) return FMC_OK;
  hipLaunchKernelGGL(philox_blocks_kernel, dim3(grid1(n_blocks, 256)), dim3(256), 0, (hipStream_t)stream, key0,
                     key1, ctr0, ctr1, ctr2, ctr3, n_blocks, out);
  return check_launch("philox blocks");
}

extern "C" int fmc_rng_normal(void* stream, uint32_t key0, uint32_t key1, int64_t n_chains, int32_t n_dim,
                              int32_t n_steps, int32_t step0, int32_t stream_id, float* out) {
  FMC_ARG(out != nullptr && n_chains >= 0 && n_dim >= 1 && n_steps >= 0 && step0 >= 0, "bad arguments");
  int64_t n = (int64_t)n_steps * n_chains * n_dim;
  if (n == 0) return FMC_OK;
  hipLaunchKernelGGL(rng_normal_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, key0, key1,
                     n_chains, n_dim, n_steps, step0, (uint32_t)stream_id, out);
  return check_launch("rng normal");
}

extern "C" int fmc_rng_log_uniform(void* stream, uint32_t key0, uint32_t key1, int64_t n_chains, int32_t n_steps,
                                   int32_t step0, int32_t stream_id, float* out) {
  FMC_ARG(out != nullptr && n_chains >= 0 && n_steps >= 0 && step0 >= 0, "bad arguments");
  int64_t n = (int64_t)n_steps * n_chains;
  if (n == 0) return FMC_OK;
  hipLaunchKernelGGL(rng_log_uniform_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, key0, key1,
                     n_chains, n_steps, step0, (uint32_t)stream_id, out);
  return check_launch("rng log uniform");
}
