// RealNVP flow kernels for gfx950 (src/flowMC/resource/model/nf_model/
// realNVP.py). Per layer a masked coupling layer whose scale and shift MLPs
// ([d, hidden, d], ReLU; common.py:68-112, 171-208) read the conditioning
// dims and drive the transformed dims:
//   s = tanh(MLP_s(c)) dt,  t = MLP_t(c) dt,  y = (x + t) e^s,  log-det = sum s.
// The model is FP32 end to end. Its GEMMs are small (d x hidden per sample and
// layer), and on CDNA4 the FP32 vector FMA rate equals the FP32 MFMA rate, so
// the products are FP32 FMA loops over LDS-resident activations (samples in
// float4 groups) with weights streamed from an L2-resident packed copy.
//
// Work mapping: a workgroup of 256 threads owns RN_TS = 32 samples for the
// whole pass (all layers), activations in LDS. The adjoint runs one layer per
// launch and writes per-tile weight-gradient partials; a second kernel sums
// them over tiles in a fixed order and scatters them into the gradient vector.
#include "fmc_common.h"

namespace fmc {

constexpr int RN_TS = 32;       // samples per workgroup
constexpr int RN_THREADS = 256;
constexpr int RN_Q = RN_TS / 4;  // float4 sample groups

struct RnvpDev {
  int d, L, h;
  float dt;
  const float* theta;
  const float* packed;
  const int32_t* layout;
};

// ---- flat parameter layout (theta): per layer, for p = scale, shift:
//   W1 [h][d], b1 [h], W2 [d][h], b2 [d]; then data_mean [d], data_cov [d][d],
//   base_mean [d], base_cov [d][d].
__host__ __device__ inline int64_t rn_mlp_size(int d, int h) { return 2LL * h * d + h + d; }
__host__ __device__ inline int64_t rn_layer_size(int d, int h) { return 2 * rn_mlp_size(d, h); }
__host__ __device__ inline int64_t rn_mlp(int d, int h, int l, int p) {
  return l * rn_layer_size(d, h) + p * rn_mlp_size(d, h);
}
__host__ __device__ inline int64_t rn_stats(int d, int h, int L) { return L * rn_layer_size(d, h); }
__host__ __device__ inline int64_t rn_theta_size(int d, int h, int L) {
  return rn_stats(d, h, L) + 2LL * d + 2LL * d * d;
}
// ---- packed copy, per (layer, p): W1T [d][h] (row k = k-th conditioning
// dim), b1 [h], b2c [d] (entry i = i-th transformed dim); the second-layer
// weights are read from theta (rows of W2 are contiguous in the hidden index).
__host__ __device__ inline int64_t rn_pk_mlp(int d, int h) { return (int64_t)d * h + h + d; }
__host__ __device__ inline int64_t rn_pk(int d, int h, int l, int p) { return (2LL * l + p) * rn_pk_mlp(d, h); }
// ---- layout: per layer [n_cond, n_trans, cond[d], trans[d]]
__host__ __device__ inline const int32_t* rn_lay(const RnvpDev& f, int l) { return f.layout + (int64_t)l * (2 + 2 * f.d); }

__host__ __device__ inline int rn_rup16(int v) { return (v + 15) & ~15; }

struct RnSmem {
  int xs, gb, cT, hT, ghT, sT, tT, ldr, glp, total;
  int ldx, d4, h4;
  __host__ __device__ RnSmem(int d, int h, bool bwd) {
    ldx = d + 1;
    d4 = (d + 3) & ~3;
    h4 = (h + 3) & ~3;
    int o = 0;
    xs = o;
    o += rn_rup16(RN_TS * ldx * 4);
    gb = o;
    o += bwd ? rn_rup16(RN_TS * ldx * 4) : 0;
    cT = o;  // [d4][TS] conditioning inputs, sample-contiguous
    o += d4 * RN_TS * 4;
    hT = o;  // [h4][TS] hidden activations
    o += h4 * RN_TS * 4;
    ghT = o;  // [h4][TS] hidden pre-activation gradients (adjoint)
    o += bwd ? h4 * RN_TS * 4 : 0;
    sT = o;  // [d4][TS] scale pre-activations / tanh / their gradients
    o += d4 * RN_TS * 4;
    tT = o;  // [d4][TS] shift outputs / their gradients
    o += d4 * RN_TS * 4;
    ldr = o;
    o += RN_TS * 4;
    glp = o;
    o += RN_TS * 4;
    total = o;
  }
};

// H[j][m] = relu(b1[j] + sum_k W1T[k][j] c[k][m]) for the 32 samples of the tile.
__device__ __forceinline__ void rn_gemm1(const float* __restrict__ W1T, const float* __restrict__ b1, int h, int dc,
                                         const float* cT, float* hT) {
  for (int o = threadIdx.x; o < h * RN_Q; o += RN_THREADS) {
    const int j = o % h, q = o / h;
    const float b = b1[j];
    float4 acc = make_float4(b, b, b, b);
    for (int k = 0; k < dc; ++k) {
      const float w = W1T[(int64_t)k * h + j];
      const float4 c = *reinterpret_cast<const float4*>(cT + k * RN_TS + 4 * q);
      acc.x = fmaf(w, c.x, acc.x);
      acc.y = fmaf(w, c.y, acc.y);
      acc.z = fmaf(w, c.z, acc.z);
      acc.w = fmaf(w, c.w, acc.w);
    }
    acc.x = fmaxf(acc.x, 0.f);
    acc.y = fmaxf(acc.y, 0.f);
    acc.z = fmaxf(acc.z, 0.f);
    acc.w = fmaxf(acc.w, 0.f);
    *reinterpret_cast<float4*>(hT + j * RN_TS + 4 * q) = acc;
  }
}

// out[i][m] = b2c[i] + sum_j W2[trans_i][j] H[j][m] for the transformed dims i < dtr.
__device__ __forceinline__ void rn_gemm2(const float* __restrict__ W2, const float* __restrict__ b2c, int h, int dtr,
                                         const int32_t* trans, const float* hT, float* outT) {
  for (int o = threadIdx.x; o < dtr * RN_Q; o += RN_THREADS) {
    const int i = o % dtr, q = o / dtr;
    const float* w = W2 + (int64_t)trans[i] * h;
    const float b = b2c[i];
    float4 acc = make_float4(b, b, b, b);
    for (int j = 0; j < h; ++j) {
      const float wj = w[j];
      const float4 v = *reinterpret_cast<const float4*>(hT + j * RN_TS + 4 * q);
      acc.x = fmaf(wj, v.x, acc.x);
      acc.y = fmaf(wj, v.y, acc.y);
      acc.z = fmaf(wj, v.z, acc.z);
      acc.w = fmaf(wj, v.w, acc.w);
    }
    *reinterpret_cast<float4*>(outT + i * RN_TS + 4 * q) = acc;
  }
}

// Conditioning inputs of layer l into cT (rows >= dc zero, up to d4).
__device__ __forceinline__ void rn_load_cond(const float* xs, int ldx, const int32_t* cond, int dc, int d4, float* cT) {
  for (int o = threadIdx.x; o < d4 * RN_TS; o += RN_THREADS) {
    const int k = o / RN_TS, m = o - k * RN_TS;
    cT[o] = (k < dc) ? xs[m * ldx + cond[k]] : 0.f;
  }
}

// ------------------------------------------------------------- packing ---
__global__ void __launch_bounds__(256) rnvp_pack_kernel(RnvpDev f, float* packed) {
  const int d = f.d, h = f.h;
  const int lp = blockIdx.y, l = lp >> 1, p = lp & 1;
  const int32_t* lay = rn_lay(f, l);
  const int dc = lay[0], dtr = lay[1];
  const int32_t* cond = lay + 2;
  const int32_t* trans = lay + 2 + d;
  const float* src = f.theta + rn_mlp(d, h, l, p);
  float* dst = packed + rn_pk(d, h, l, p);
  const int64_t n = rn_pk_mlp(d, h);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    float v;
    if (e < (int64_t)d * h) {  // W1T[k][j] = W1[j][cond_k]
      const int k = (int)(e / h), j = (int)(e - (int64_t)k * h);
      v = (k < dc) ? src[(int64_t)j * d + cond[k]] : 0.f;
    } else if (e < (int64_t)d * h + h) {
      v = src[(int64_t)h * d + (e - (int64_t)d * h)];
    } else {
      const int i = (int)(e - (int64_t)d * h - h);
      v = (i < dtr) ? src[2LL * h * d + h + trans[i]] : 0.f;
    }
    dst[e] = v;
  }
}

// -------------------------------------------------------------- passes ---
struct RnPassArgs {
  int dir;    // 0 forward (realNVP.py:163-180), 1 inverse (:182-198)
  int flags;  // 1 standardize, 2 base log-prob, 4 sample base noise
  const float* x_in;
  int64_t n;
  uint32_t key0, key1;
  int64_t sample0;
  float* x_out;
  float* logdet;
  float* saved;
  const float* g_logp;
  float* gy_out;
};

enum { RN_STD = 1, RN_BASE = 2, RN_SAMPLE = 4 };

__global__ void __launch_bounds__(RN_THREADS) rnvp_pass_kernel(RnvpDev f, RnPassArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int d = f.d, h = f.h;
  const RnSmem sm(d, h, false);
  float* xs = reinterpret_cast<float*>(smem + sm.xs);
  float* cT = reinterpret_cast<float*>(smem + sm.cT);
  float* hT = reinterpret_cast<float*>(smem + sm.hT);
  float* sT = reinterpret_cast<float*>(smem + sm.sT);
  float* tT = reinterpret_cast<float*>(smem + sm.tT);
  float* ldr = reinterpret_cast<float*>(smem + sm.ldr);
  const int tid = threadIdx.x, ldx = sm.ldx;
  const int64_t m0 = (int64_t)blockIdx.x * RN_TS;
  const float* th = f.theta;
  const int64_t st = rn_stats(d, h, f.L);
  const float* dmean = th + st;
  const float* dcov = dmean + d;
  const float* bmean = dcov + (int64_t)d * d;
  const float* bcov = bmean + d;

  for (int o = tid; o < RN_TS * d; o += RN_THREADS) {
    const int m = o / d, j = o - m * d;
    const int64_t row = m0 + m;
    float v = 0.f;
    if (row < a.n) {
      if ((a.flags & RN_SAMPLE) && a.x_in == nullptr) {
        const float4 z = philox_normal4(a.key0, a.key1, (uint32_t)(j >> 2), (uint32_t)(a.sample0 + row), 0u,
                                        RNG_FLOW_BASE);
        const int e = j & 3;
        v = (e == 0) ? z.x : (e == 1) ? z.y : (e == 2) ? z.z : z.w;
      } else {
        v = a.x_in[row * d + j];
      }
      if (a.dir == 0 && (a.flags & RN_STD)) {
        v = (v - dmean[j]) / sqrtf(dcov[(int64_t)j * d + j]);
      } else if (a.dir != 0 && (a.flags & RN_SAMPLE)) {
        v = bmean[j] + sqrtf(bcov[(int64_t)j * d + j]) * v;
      }
    }
    xs[m * ldx + j] = v;
  }
  for (int m = tid; m < RN_TS; m += RN_THREADS) ldr[m] = 0.f;
  __syncthreads();

  for (int step = 0; step < f.L; ++step) {
    const int l = (a.dir == 0) ? step : f.L - 1 - step;
    if (a.saved) {
      for (int o = tid; o < RN_TS * d; o += RN_THREADS) {
        const int m = o / d, j = o - m * d;
        if (m0 + m < a.n) a.saved[((int64_t)l * a.n + m0 + m) * d + j] = xs[m * ldx + j];
      }
    }
    const int32_t* lay = rn_lay(f, l);
    const int dc = lay[0], dtr = lay[1];
    const int32_t* cond = lay + 2;
    const int32_t* trans = lay + 2 + d;
    rn_load_cond(xs, ldx, cond, dc, sm.d4, cT);
    __syncthreads();
    for (int p = 0; p < 2; ++p) {
      const float* pk = f.packed + rn_pk(d, h, l, p);
      rn_gemm1(pk, pk + (int64_t)d * h, h, dc, cT, hT);
      __syncthreads();
      rn_gemm2(th + rn_mlp(d, h, l, p) + (int64_t)h * d + h, pk + (int64_t)d * h + h, h, dtr, trans, hT,
               p == 0 ? sT : tT);
      __syncthreads();
    }
    // coupling on the transformed dims; sT keeps s for the log-det sum
    for (int o = tid; o < dtr * RN_TS; o += RN_THREADS) {
      const int i = o / RN_TS, m = o - i * RN_TS;
      const int j = trans[i];
      const float s = tanhf(sT[o]) * f.dt;
      const float t = tT[o] * f.dt;
      const float x = xs[m * ldx + j];
      xs[m * ldx + j] = (a.dir == 0) ? (x + t) * expf(s) : x * expf(-s) - t;
      sT[o] = s;
    }
    __syncthreads();
    for (int m = tid; m < RN_TS; m += RN_THREADS) {
      float s = 0.f;
      for (int i = 0; i < dtr; ++i) s += sT[i * RN_TS + m];
      ldr[m] += (a.dir == 0) ? s : -s;
    }
    __syncthreads();
  }
  // ---- outputs
  for (int m = tid; m < RN_TS; m += RN_THREADS) {
    const int64_t row = m0 + m;
    if (row >= a.n) continue;
    float ld = ldr[m];
    if (a.dir == 0 && (a.flags & RN_BASE)) {
      // log N(y; 0, I) (realNVP.py:211-213: the standard normal, not base_dist)
      float q = 0.f;
      for (int j = 0; j < d; ++j) {
        const float y = xs[m * ldx + j];
        q += y * y;
        if (a.gy_out) a.gy_out[row * d + j] = -a.g_logp[row] * y;
      }
      ld += -0.5f * q - 0.91893853320467274f * (float)d;
    }
    if (a.logdet) a.logdet[row] = ld;
  }
  if (a.x_out) {
    for (int o = tid; o < RN_TS * d; o += RN_THREADS) {
      const int m = o / d, j = o - m * d;
      const int64_t row = m0 + m;
      if (row >= a.n) continue;
      float v = xs[m * ldx + j];
      if (a.dir != 0 && (a.flags & RN_STD)) v = v * sqrtf(dcov[(int64_t)j * d + j]) + dmean[j];
      a.x_out[row * d + j] = v;
    }
  }
}

// ------------------------------------------------------------- adjoint ---
// One layer in the forward direction: recompute s, t; the coupling's
// reverse sweep; per MLP the hidden gradient (W2^T g), its ReLU mask, the
// tile's weight-gradient partials, and W1^T g into the conditioning dims.
// partial layout per tile (one layer): for p = scale, shift:
//   dW1 [h][d] (column k = k-th conditioning dim), db1 [h],
//   dW2 [d][h] (row i = i-th transformed dim), db2 [d].
struct RnBwdArgs {
  int layer;
  const float* x_in;  // [n, d] layer input
  int64_t n;
  const float* g_logp;
  const float* gy;  // [n, d] dL/d(layer output)
  float* gx;        // [n, d] dL/d(layer input)
  float* part;      // [n_tiles, layer_size]
};

// out[a][b] = sum_m A[a][m] B[b][m] over the tile's 32 samples, 4x4 register
// tiles; A has na rows (storage padded to na4), B nb rows (padded to nb4).
__device__ __forceinline__ void rn_outer(const float* A, int na, const float* B, int nb, float* out, int ldo) {
  const int na4 = (na + 3) >> 2, nb4 = (nb + 3) >> 2;
  for (int o = threadIdx.x; o < na4 * nb4; o += RN_THREADS) {
    const int ta = o / nb4, tb = o - ta * nb4;
    float acc[4][4] = {};
#pragma unroll 1
    for (int q = 0; q < RN_Q; ++q) {
      float4 av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        av[u] = *reinterpret_cast<const float4*>(A + (4 * ta + u) * RN_TS + 4 * q);
        bv[u] = *reinterpret_cast<const float4*>(B + (4 * tb + u) * RN_TS + 4 * q);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          acc[u][v] += av[u].x * bv[v].x + av[u].y * bv[v].y + av[u].z * bv[v].z + av[u].w * bv[v].w;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ia = 4 * ta + u, ib = 4 * tb + v;
        if (ia < na && ib < nb) out[(int64_t)ia * ldo + ib] = acc[u][v];
      }
  }
}

__device__ __forceinline__ void rn_rowsum(const float* A, int na, float* out) {
  for (int i = threadIdx.x; i < na; i += RN_THREADS) {
    float s = 0.f;
    for (int m = 0; m < RN_TS; ++m) s += A[i * RN_TS + m];
    out[i] = s;
  }
}

__global__ void __launch_bounds__(RN_THREADS) rnvp_bwd_kernel(RnvpDev f, RnBwdArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int d = f.d, h = f.h, l = a.layer;
  const RnSmem sm(d, h, true);
  float* xs = reinterpret_cast<float*>(smem + sm.xs);
  float* gb = reinterpret_cast<float*>(smem + sm.gb);
  float* cT = reinterpret_cast<float*>(smem + sm.cT);
  float* hT = reinterpret_cast<float*>(smem + sm.hT);
  float* ghT = reinterpret_cast<float*>(smem + sm.ghT);
  float* sT = reinterpret_cast<float*>(smem + sm.sT);
  float* tT = reinterpret_cast<float*>(smem + sm.tT);
  float* glp = reinterpret_cast<float*>(smem + sm.glp);
  const int tid = threadIdx.x, ldx = sm.ldx;
  const int64_t m0 = (int64_t)blockIdx.x * RN_TS;
  const int32_t* lay = rn_lay(f, l);
  const int dc = lay[0], dtr = lay[1];
  const int32_t* cond = lay + 2;
  const int32_t* trans = lay + 2 + d;

  for (int o = tid; o < RN_TS * d; o += RN_THREADS) {
    const int m = o / d, j = o - m * d;
    const int64_t row = m0 + m;
    const bool ok = row < a.n;
    xs[m * ldx + j] = ok ? a.x_in[row * d + j] : 0.f;
    gb[m * ldx + j] = ok ? a.gy[row * d + j] : 0.f;
  }
  for (int m = tid; m < RN_TS; m += RN_THREADS) glp[m] = (m0 + m < a.n) ? a.g_logp[m0 + m] : 0.f;
  // zero the padding rows of the transposed buffers (read by the 4x4 tiles)
  for (int o = tid; o < (sm.d4 - dtr) * RN_TS; o += RN_THREADS) {
    sT[dtr * RN_TS + o] = 0.f;
    tT[dtr * RN_TS + o] = 0.f;
  }
  for (int o = tid; o < (sm.h4 - h) * RN_TS; o += RN_THREADS) {
    hT[h * RN_TS + o] = 0.f;
    ghT[h * RN_TS + o] = 0.f;
  }
  __syncthreads();
  rn_load_cond(xs, ldx, cond, dc, sm.d4, cT);
  __syncthreads();
  // ---- recompute scale / shift pre-activations
  for (int p = 0; p < 2; ++p) {
    const float* pk = f.packed + rn_pk(d, h, l, p);
    rn_gemm1(pk, pk + (int64_t)d * h, h, dc, cT, hT);
    __syncthreads();
    rn_gemm2(f.theta + rn_mlp(d, h, l, p) + (int64_t)h * d + h, pk + (int64_t)d * h + h, h, dtr, trans, hT,
             p == 0 ? sT : tT);
    __syncthreads();
  }
  // ---- coupling reverse sweep: y = (x + t) e^s, log-det += s
  //   g_x = g e^s (transformed dim), g_t = g e^s, g_s = g y + g_logp;
  //   g_spre = g_s dt (1 - tanh^2), g_tpre = g_t dt (in place of sT / tT)
  for (int o = tid; o < dtr * RN_TS; o += RN_THREADS) {
    const int i = o / RN_TS, m = o - i * RN_TS;
    const int j = trans[i];
    const float tn = tanhf(sT[o]);
    const float s = tn * f.dt;
    const float t = tT[o] * f.dt;
    const float e = expf(s);
    const float x = xs[m * ldx + j];
    const float y = (x + t) * e;
    const float g = gb[m * ldx + j];
    const float ge = g * e;
    gb[m * ldx + j] = ge;
    sT[o] = (g * y + glp[m]) * f.dt * (1.f - tn * tn);
    tT[o] = ge * f.dt;
  }
  __syncthreads();
  float* part = a.part + (int64_t)blockIdx.x * rn_layer_size(d, h);
  for (int p = 0; p < 2; ++p) {
    const float* pk = f.packed + rn_pk(d, h, l, p);
    const float* W1 = f.theta + rn_mlp(d, h, l, p);
    const float* W2 = W1 + (int64_t)h * d + h;
    const float* goT = (p == 0) ? sT : tT;
    float* pp = part + p * rn_mlp_size(d, h);
    rn_gemm1(pk, pk + (int64_t)d * h, h, dc, cT, hT);
    __syncthreads();
    // g_H[j][m] = sum_i W2[trans_i][j] g_out[i][m], masked by ReLU' (H > 0)
    for (int o = tid; o < h * RN_Q; o += RN_THREADS) {
      const int j = o % h, q = o / h;
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int i = 0; i < dtr; ++i) {
        const float w = W2[(int64_t)trans[i] * h + j];
        const float4 g = *reinterpret_cast<const float4*>(goT + i * RN_TS + 4 * q);
        acc.x = fmaf(w, g.x, acc.x);
        acc.y = fmaf(w, g.y, acc.y);
        acc.z = fmaf(w, g.z, acc.z);
        acc.w = fmaf(w, g.w, acc.w);
      }
      const float4 hv = *reinterpret_cast<const float4*>(hT + j * RN_TS + 4 * q);
      acc.x = hv.x > 0.f ? acc.x : 0.f;
      acc.y = hv.y > 0.f ? acc.y : 0.f;
      acc.z = hv.z > 0.f ? acc.z : 0.f;
      acc.w = hv.w > 0.f ? acc.w : 0.f;
      *reinterpret_cast<float4*>(ghT + j * RN_TS + 4 * q) = acc;
    }
    __syncthreads();
    // tile partials: dW1[j][k] = sum_m gH[j][m] c[k][m]; db1; dW2[i][j] = sum_m g_out[i][m] H[j][m]; db2
    rn_outer(ghT, h, cT, dc, pp, d);
    rn_rowsum(ghT, h, pp + (int64_t)h * d);
    rn_outer(goT, dtr, hT, h, pp + (int64_t)h * d + h, h);
    rn_rowsum(goT, dtr, pp + 2LL * h * d + h);
    // g_c[m][k] += sum_j W1[j][cond_k] gH[j][m]   (one thread per (k, 4 samples))
    for (int o = tid; o < dc * RN_Q; o += RN_THREADS) {
      const int k = o % dc, q = o / dc;
      const float* w = W1 + cond[k];
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int j = 0; j < h; ++j) {
        const float wj = w[(int64_t)j * d];
        const float4 g = *reinterpret_cast<const float4*>(ghT + j * RN_TS + 4 * q);
        acc.x = fmaf(wj, g.x, acc.x);
        acc.y = fmaf(wj, g.y, acc.y);
        acc.z = fmaf(wj, g.z, acc.z);
        acc.w = fmaf(wj, g.w, acc.w);
      }
      const int jd = cond[k];
      gb[(4 * q + 0) * ldx + jd] += acc.x;
      gb[(4 * q + 1) * ldx + jd] += acc.y;
      gb[(4 * q + 2) * ldx + jd] += acc.z;
      gb[(4 * q + 3) * ldx + jd] += acc.w;
    }
    __syncthreads();
  }
  for (int o = tid; o < RN_TS * d; o += RN_THREADS) {
    const int m = o / d, j = o - m * d;
    const int64_t row = m0 + m;
    if (row < a.n) a.gx[row * d + j] = gb[m * ldx + j];
  }
}

// Sum the tile partials of one layer in tile order and scatter them to the
// reference-shaped leaves of the gradient vector.
__global__ void __launch_bounds__(256) rnvp_wgrad_kernel(RnvpDev f, int l, const float* __restrict__ part,
                                                         int64_t n_tiles, float* __restrict__ grad) {
  const int d = f.d, h = f.h;
  const int64_t LS = rn_layer_size(d, h), MS = rn_mlp_size(d, h);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= LS) return;
  const int32_t* lay = rn_lay(f, l);
  const int dc = lay[0], dtr = lay[1];
  const int32_t* cond = lay + 2;
  const int32_t* trans = lay + 2 + d;
  const int p = (int)(e / MS);
  const int64_t r = e - p * MS;
  int64_t dst = -1;
  if (r < (int64_t)h * d) {
    const int j = (int)(r / d), k = (int)(r - (int64_t)j * d);
    if (k < dc) dst = (int64_t)j * d + cond[k];
  } else if (r < (int64_t)h * d + h) {
    dst = r;
  } else if (r < 2LL * h * d + h) {
    const int64_t rr = r - (int64_t)h * d - h;
    const int i = (int)(rr / h), j = (int)(rr - (int64_t)i * h);
    if (i < dtr) dst = (int64_t)h * d + h + (int64_t)trans[i] * h + j;
  } else {
    const int i = (int)(r - 2LL * h * d - h);
    if (i < dtr) dst = 2LL * h * d + h + trans[i];
  }
  if (dst < 0) return;
  float s = 0.f;
  for (int64_t t = 0; t < n_tiles; ++t) s += part[t * LS + e];
  grad[rn_mlp(d, h, l, p) + dst] = s;
}

static int check_rnvp(const fmc_rnvp* r, RnvpDev& f, bool need_packed) {
  FMC_ARG(r != nullptr, "RealNVP description is NULL");
  FMC_ARG(r->n_dim >= 1 && r->n_layers >= 1 && r->n_hidden >= 1, "n_dim, n_layers and n_hidden must be >= 1");
  FMC_ARG(r->theta != nullptr && r->layout != nullptr, "theta or layout is NULL");
  FMC_ARG(!need_packed || r->packed != nullptr, "packed buffer is NULL");
  const RnSmem sm(r->n_dim, r->n_hidden, true);
  FMC_ARG(sm.total <= 160 * 1024, "n_dim %d / n_hidden %d exceed the LDS budget of the RealNVP kernels",
          r->n_dim, r->n_hidden);
  f = RnvpDev{r->n_dim, r->n_layers, r->n_hidden, r->dt, r->theta, r->packed, r->layout};
  return FMC_OK;
}

template <class K>
static int rn_set_lds(K kernel, size_t shm) {
  const hipError_t e =
      hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) {
    set_error(FMC_ERR_HIP, "hipFuncSetAttribute(%zu bytes of LDS): %s", shm, hipGetErrorString(e));
    return FMC_ERR_HIP;
  }
  return FMC_OK;
}

}  // namespace fmc

using namespace fmc;

extern "C" int fmc_rnvp_sizes(const fmc_rnvp* r, int64_t* theta_size, int64_t* packed_size,
                              int64_t* partial_per_tile, int32_t* tile) {
  RnvpDev f;
  int rc = check_rnvp(r, f, false);
  if (rc) return rc;
  if (theta_size) *theta_size = rn_theta_size(f.d, f.h, f.L);
  if (packed_size) *packed_size = 2LL * f.L * rn_pk_mlp(f.d, f.h);
  if (partial_per_tile) *partial_per_tile = rn_layer_size(f.d, f.h);
  if (tile) *tile = RN_TS;
  return FMC_OK;
}

extern "C" int fmc_rnvp_pack(void* stream, const fmc_rnvp* r) {
  RnvpDev f;
  int rc = check_rnvp(r, f, true);
  if (rc) return rc;
  const int64_t n = rn_pk_mlp(f.d, f.h);
  const unsigned gx = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  hipLaunchKernelGGL(rnvp_pack_kernel, dim3(gx, 2 * f.L), dim3(256), 0, (hipStream_t)stream, f, r->packed);
  return check_launch("realnvp pack");
}

extern "C" int fmc_rnvp_pass(void* stream, const fmc_rnvp* r, int32_t dir, int32_t flags, const float* x_in,
                             int64_t n, uint32_t key0, uint32_t key1, int64_t sample0, float* x_out, float* logdet,
                             float* saved, const float* g_logp, float* gy_out) {
  RnvpDev f;
  int rc = check_rnvp(r, f, true);
  if (rc) return rc;
  FMC_ARG(dir == 0 || dir == 1, "dir must be 0 (forward) or 1 (inverse)");
  FMC_ARG(n >= 0, "n < 0");
  FMC_ARG(x_in != nullptr || (flags & RN_SAMPLE), "x_in is NULL and base sampling is off");
  FMC_ARG(!(flags & RN_SAMPLE) || dir == 1, "base sampling needs the inverse direction");
  FMC_ARG(!(flags & RN_BASE) || dir == 0, "the base log-prob is added in the forward direction only");
  FMC_ARG(saved == nullptr || dir == 0, "saved layer inputs need the forward direction");
  FMC_ARG(gy_out == nullptr || (g_logp != nullptr && (flags & RN_BASE)), "gy_out needs g_logp and BASE_LOGP");
  if (n == 0) return FMC_OK;
  const RnSmem sm(f.d, f.h, false);
  if ((rc = rn_set_lds(rnvp_pass_kernel, (size_t)sm.total))) return rc;
  RnPassArgs a{dir, flags, x_in, n, key0, key1, sample0, x_out, logdet, saved, g_logp, gy_out};
  const int64_t grid = (n + RN_TS - 1) / RN_TS;
  FMC_ARG(grid <= 0x7fffffffLL, "too many samples for one launch");
  hipLaunchKernelGGL(rnvp_pass_kernel, dim3((unsigned)grid), dim3(RN_THREADS), (size_t)sm.total,
                     (hipStream_t)stream, f, a);
  return check_launch("realnvp pass");
}

extern "C" int fmc_rnvp_backward_layer(void* stream, const fmc_rnvp* r, int32_t layer, const float* x_in, int64_t n,
                                       const float* g_logp, const float* gy, float* gx, float* partials,
                                       int64_t n_partial_tiles) {
  RnvpDev f;
  int rc = check_rnvp(r, f, true);
  if (rc) return rc;
  FMC_ARG(layer >= 0 && layer < f.L, "layer %d out of range", layer);
  FMC_ARG(n >= 0, "n < 0");
  FMC_ARG(x_in && g_logp && gy && gx && partials, "NULL pointer");
  FMC_ARG(gy != gx, "gx must not alias gy");
  const int64_t tiles = (n + RN_TS - 1) / RN_TS;
  FMC_ARG(n_partial_tiles >= tiles, "partials hold %lld tiles, %lld needed", (long long)n_partial_tiles,
          (long long)tiles);
  if (n == 0) return FMC_OK;
  const RnSmem sm(f.d, f.h, true);
  if ((rc = rn_set_lds(rnvp_bwd_kernel, (size_t)sm.total))) return rc;
  RnBwdArgs a{layer, x_in, n, g_logp, gy, gx, partials};
  hipLaunchKernelGGL(rnvp_bwd_kernel, dim3((unsigned)tiles), dim3(RN_THREADS), (size_t)sm.total,
                     (hipStream_t)stream, f, a);
  return check_launch("realnvp adjoint");
}

extern "C" int fmc_rnvp_weight_grad(void* stream, const fmc_rnvp* r, int32_t layer, const float* partials,
                                    int64_t n_tiles, float* grad) {
  RnvpDev f;
  int rc = check_rnvp(r, f, false);
  if (rc) return rc;
  FMC_ARG(layer >= 0 && layer < f.L, "layer %d out of range", layer);
  FMC_ARG(partials && grad, "NULL pointer");
  FMC_ARG(n_tiles >= 0, "n_tiles < 0");
  const int64_t LS = rn_layer_size(f.d, f.h);
  hipLaunchKernelGGL(rnvp_wgrad_kernel, dim3((unsigned)((LS + 255) / 256)), dim3(256), 0, (hipStream_t)stream, f,
                     layer, partials, n_tiles, grad);
  return check_launch("realnvp weight gradient");
}
