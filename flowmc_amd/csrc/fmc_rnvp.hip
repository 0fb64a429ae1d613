// RealNVP flow kernels for gfx950 (src/flowMC/resource/model/nf_model/
// realNVP.py). Per layer a masked coupling layer whose scale and shift MLPs
// ([d, hidden, d], ReLU; common.py:68-112, 171-208) read the conditioning
// dims and drive the transformed dims:
//   s = tanh(MLP_s(c)) dt,  t = MLP_t(c) dt,  y = (x + t) e^s,  log-det = sum s.
// The model is FP32 end to end. Its GEMMs are small (d x hidden per sample and
// layer), and on CDNA4 the FP32 vector FMA rate equals the FP32 MFMA rate, so
// the products are FP32 FMA register tiles (4 outputs x 4 samples per thread)
// over LDS-resident activations, with weights read as aligned float4 rows of
// an L2-resident packed copy. Products with few outputs (the d_t scale / shift
// rows, the d_c conditioning gradients) split their reduction over the lanes
// of a wavefront and combine them with xor shuffles.
//
// Work mapping: a workgroup of 256 threads owns RN_TS = 16 samples for the
// whole pass (all layers), activations in LDS (~42 KB at d=32, hidden 256, so
// three workgroups share a CU). The adjoint runs one layer per launch and
// writes per-tile weight-gradient partials; a second kernel sums them over
// tiles in a fixed order and scatters them into the gradient vector.
#include "fmc_common.h"

namespace fmc {

constexpr int RN_TS = 16;       // samples per workgroup
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
// ---- packed copy, per (layer, p), with d4 / h4 = d / h rounded up to 4 and
// zero padding: W1T [d4][h4] (W1T[k][j] = W1[j][cond_k]), W1c [h4][d4] (its
// transpose), W2T [h4][d4] (W2T[j][i] = W2[trans_i][j]), W2c [d4][h4] (its
// transpose), b1 [h4], b2c [d4] (b2c[i] = b2[trans_i]). Every row starts on a
// 16-byte boundary.
__host__ __device__ inline int rn_r4(int v) { return (v + 3) & ~3; }
__host__ __device__ inline int64_t rn_pk_mlp(int d, int h) {
  return 4LL * rn_r4(d) * rn_r4(h) + rn_r4(h) + rn_r4(d);
}
__host__ __device__ inline int64_t rn_pk(int d, int h, int l, int p) { return (2LL * l + p) * rn_pk_mlp(d, h); }
struct RnPk {
  const float *W1T, *W1c, *W2T, *W2c, *b1, *b2c;
  __device__ RnPk(const float* base, int d4, int h4) {
    const int64_t s = (int64_t)d4 * h4;
    W1T = base;
    W1c = base + s;
    W2T = base + 2 * s;
    W2c = base + 3 * s;
    b1 = base + 4 * s;
    b2c = b1 + h4;
  }
};
// ---- layout: per layer [n_cond, n_trans, cond[d], trans[d]]
__host__ __device__ inline const int32_t* rn_lay(const RnvpDev& f, int l) { return f.layout + (int64_t)l * (2 + 2 * f.d); }

__host__ __device__ inline int rn_rup16(int v) { return (v + 15) & ~15; }

struct RnSmem {
  int xs, gb, cT, hT, ghT, sT, tT, gcT, ldr, glp, total;
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
    gcT = o;  // [d4][TS] conditioning-input gradients of one MLP (adjoint)
    o += bwd ? d4 * RN_TS * 4 : 0;
    ldr = o;
    o += RN_TS * 4;
    glp = o;
    o += RN_TS * 4;
    total = o;
  }
};

#define RN_FMA16(acc, w, c)                    \
  do {                                         \
    acc[0][0] = fmaf(w.x, c.x, acc[0][0]);     \
    acc[0][1] = fmaf(w.x, c.y, acc[0][1]);     \
    acc[0][2] = fmaf(w.x, c.z, acc[0][2]);     \
    acc[0][3] = fmaf(w.x, c.w, acc[0][3]);     \
    acc[1][0] = fmaf(w.y, c.x, acc[1][0]);     \
    acc[1][1] = fmaf(w.y, c.y, acc[1][1]);     \
    acc[1][2] = fmaf(w.y, c.z, acc[1][2]);     \
    acc[1][3] = fmaf(w.y, c.w, acc[1][3]);     \
    acc[2][0] = fmaf(w.z, c.x, acc[2][0]);     \
    acc[2][1] = fmaf(w.z, c.y, acc[2][1]);     \
    acc[2][2] = fmaf(w.z, c.z, acc[2][2]);     \
    acc[2][3] = fmaf(w.z, c.w, acc[2][3]);     \
    acc[3][0] = fmaf(w.w, c.x, acc[3][0]);     \
    acc[3][1] = fmaf(w.w, c.y, acc[3][1]);     \
    acc[3][2] = fmaf(w.w, c.z, acc[3][2]);     \
    acc[3][3] = fmaf(w.w, c.w, acc[3][3]);     \
  } while (0)

// Wide product, h4 output rows: out[r][m] = sum_k W[k][r] in[k][m] (W rows of
// length h4), one 4-row x 4-sample tile per thread. MODE 0: out = relu(bias +
// .); MODE 1: out = . where hmask > 0, else 0 (ReLU derivative). Padding rows
// come out zero (zero weights and bias).
template <int MODE>
__device__ __forceinline__ void rn_wide(const float* __restrict__ W, const float* __restrict__ bias, int K, int h4,
                                        const float* inT, float* outT, const float* hmask) {
  const int tr = h4 >> 2;
  for (int o = threadIdx.x; o < tr * RN_Q; o += RN_THREADS) {
    const int jq = o % tr, q = o / tr;
    float acc[4][4];
    const float4 b = (MODE == 0) ? *reinterpret_cast<const float4*>(bias + 4 * jq) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[0][v] = b.x;
      acc[1][v] = b.y;
      acc[2][v] = b.z;
      acc[3][v] = b.w;
    }
    for (int k = 0; k < K; ++k) {
      const float4 w = *reinterpret_cast<const float4*>(W + (int64_t)k * h4 + 4 * jq);
      const float4 c = *reinterpret_cast<const float4*>(inT + k * RN_TS + 4 * q);
      RN_FMA16(acc, w, c);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      float4 r = make_float4(acc[u][0], acc[u][1], acc[u][2], acc[u][3]);
      const int row = 4 * jq + u;
      if (MODE == 0) {
        r.x = fmaxf(r.x, 0.f);
        r.y = fmaxf(r.y, 0.f);
        r.z = fmaxf(r.z, 0.f);
        r.w = fmaxf(r.w, 0.f);
      } else {
        const float4 hv = *reinterpret_cast<const float4*>(hmask + row * RN_TS + 4 * q);
        r.x = hv.x > 0.f ? r.x : 0.f;
        r.y = hv.y > 0.f ? r.y : 0.f;
        r.z = hv.z > 0.f ? r.z : 0.f;
        r.w = hv.w > 0.f ? r.w : 0.f;
      }
      *reinterpret_cast<float4*>(outT + row * RN_TS + 4 * q) = r;
    }
  }
}

// Narrow product, n output rows (written up to n rounded to 4; the padding
// rows come out zero): out[r][m] = bias[r] + sum_{j < K} W[j][r] in[j][m] (W
// rows of length ldw). The reduction over j is split over S lanes of one
// wavefront (S a power of two, tiles * S <= 256) and combined with xor
// shuffles in a fixed order.
__device__ __forceinline__ void rn_narrow(const float* __restrict__ W, int ldw, const float* __restrict__ bias, int K,
                                          int n, const float* inT, float* outT) {
  const int tr = (n + 3) >> 2;
  const int tiles = tr * RN_Q;
  int S = 1;
  while (S < 64 && tiles * S * 2 <= RN_THREADS) S <<= 1;
  for (int o = threadIdx.x; o < tiles * S; o += RN_THREADS) {
    const int tile = o / S, s = o - tile * S;
    const int iq = tile % tr, q = tile / tr;
    float acc[4][4] = {};
    for (int j = s; j < K; j += S) {
      const float4 w = *reinterpret_cast<const float4*>(W + (int64_t)j * ldw + 4 * iq);
      const float4 c = *reinterpret_cast<const float4*>(inT + j * RN_TS + 4 * q);
      RN_FMA16(acc, w, c);
    }
    for (int off = S >> 1; off > 0; off >>= 1)
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += __shfl_xor(acc[u][v], off, 64);
    if (s == 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int row = 4 * iq + u;
        const float b = bias ? bias[row] : 0.f;
        *reinterpret_cast<float4*>(outT + row * RN_TS + 4 * q) =
            make_float4(acc[u][0] + b, acc[u][1] + b, acc[u][2] + b, acc[u][3] + b);
      }
    }
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
  const int d = f.d, h = f.h, d4 = rn_r4(d), h4 = rn_r4(h);
  const int lp = blockIdx.y, l = lp >> 1, p = lp & 1;
  const int32_t* lay = rn_lay(f, l);
  const int dc = lay[0], dtr = lay[1];
  const int32_t* cond = lay + 2;
  const int32_t* trans = lay + 2 + d;
  const float* W1 = f.theta + rn_mlp(d, h, l, p);
  const float* b1 = W1 + (int64_t)h * d;
  const float* W2 = b1 + h;
  const float* b2 = W2 + (int64_t)d * h;
  float* dst = packed + rn_pk(d, h, l, p);
  const int64_t sq = (int64_t)d4 * h4, n = rn_pk_mlp(d, h);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    float v = 0.f;
    const int region = (e < 4 * sq) ? (int)(e / sq) : 4;
    const int64_t r = e - (int64_t)region * sq;
    if (region == 0) {  // W1T[k][j]
      const int k = (int)(r / h4), j = (int)(r - (int64_t)k * h4);
      if (k < dc && j < h) v = W1[(int64_t)j * d + cond[k]];
    } else if (region == 1) {  // W1c[j][k]
      const int j = (int)(r / d4), k = (int)(r - (int64_t)j * d4);
      if (k < dc && j < h) v = W1[(int64_t)j * d + cond[k]];
    } else if (region == 2) {  // W2T[j][i]
      const int j = (int)(r / d4), i = (int)(r - (int64_t)j * d4);
      if (i < dtr && j < h) v = W2[(int64_t)trans[i] * h + j];
    } else if (region == 3) {  // W2c[i][j]
      const int i = (int)(r / h4), j = (int)(r - (int64_t)i * h4);
      if (i < dtr && j < h) v = W2[(int64_t)trans[i] * h + j];
    } else if (r < h4) {  // b1
      if (r < h) v = b1[r];
    } else {  // b2c
      const int i = (int)(r - h4);
      if (i < dtr) v = b2[trans[i]];
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
      const RnPk pk(f.packed + rn_pk(d, h, l, p), sm.d4, sm.h4);
      rn_wide<0>(pk.W1T, pk.b1, dc, sm.h4, cT, hT, nullptr);
      __syncthreads();
      rn_narrow(pk.W2T, sm.d4, pk.b2c, h, dtr, hT, p == 0 ? sT : tT);
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
  float* gcT = reinterpret_cast<float*>(smem + sm.gcT);
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
  __syncthreads();
  rn_load_cond(xs, ldx, cond, dc, sm.d4, cT);
  __syncthreads();
  // ---- recompute scale / shift pre-activations
  for (int p = 0; p < 2; ++p) {
    const RnPk pk(f.packed + rn_pk(d, h, l, p), sm.d4, sm.h4);
    rn_wide<0>(pk.W1T, pk.b1, dc, sm.h4, cT, hT, nullptr);
    __syncthreads();
    rn_narrow(pk.W2T, sm.d4, pk.b2c, h, dtr, hT, p == 0 ? sT : tT);
    __syncthreads();
  }
  // ---- coupling reverse sweep: y = (x + t) e^s, log-det += s
  //   g_x = g e^s (transformed dim), g_t = g e^s, g_s = g y + g_logp;
  //   g_spre = g_s dt (1 - tanh^2), g_tpre = g_t dt (in place of sT / tT;
  //   their padding rows stay zero)
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
    const RnPk pk(f.packed + rn_pk(d, h, l, p), sm.d4, sm.h4);
    const float* goT = (p == 0) ? sT : tT;
    float* pp = part + p * rn_mlp_size(d, h);
    rn_wide<0>(pk.W1T, pk.b1, dc, sm.h4, cT, hT, nullptr);
    __syncthreads();
    // g_H[j][m] = sum_i W2[trans_i][j] g_out[i][m], masked by ReLU' (H > 0)
    rn_wide<1>(pk.W2c, nullptr, dtr, sm.h4, goT, ghT, hT);
    __syncthreads();
    // tile partials: dW1[j][k] = sum_m gH[j][m] c[k][m]; db1; dW2[i][j] = sum_m g_out[i][m] H[j][m]; db2
    rn_outer(ghT, h, cT, dc, pp, d);
    rn_rowsum(ghT, h, pp + (int64_t)h * d);
    rn_outer(goT, dtr, hT, h, pp + (int64_t)h * d + h, h);
    rn_rowsum(goT, dtr, pp + 2LL * h * d + h);
    // g_c[k][m] = sum_j W1[j][cond_k] gH[j][m], added to the conditioning dims
    rn_narrow(pk.W1c, sm.d4, nullptr, h, dc, ghT, gcT);
    __syncthreads();
    for (int o = tid; o < dc * RN_TS; o += RN_THREADS) {
      const int k = o / RN_TS, m = o - k * RN_TS;
      gb[m * ldx + cond[k]] += gcT[o];
    }
    __syncthreads();
  }
  for (int o = tid; o < RN_TS * d; o += RN_THREADS) {
    const int m = o / d, j = o - m * d;
    const int64_t row = m0 + m;
    if (row < a.n) a.gx[row * d + j] = gb[m * ldx + j];
  }
}

// Sum of the tile partials of one layer, in two fixed-order stages so that
// enough lanes stream the partial buffer: stage 1 sums chunk c of the tiles
// (tiles c, c + C, c + 2C, ...) into row n_tiles + c of the buffer; stage 2
// sums the C chunk rows in order and scatters the result to the
// reference-shaped leaves of the gradient vector.
constexpr int RN_RED_CHUNKS = 16;

__host__ __device__ inline int rn_chunks(int64_t n_tiles) {
  const int64_t c = (n_tiles + 15) / 16;
  return (int)(c < 1 ? 1 : (c > RN_RED_CHUNKS ? RN_RED_CHUNKS : c));
}

__global__ void __launch_bounds__(256) rnvp_wgrad_stage1(int64_t LS, float* part, int64_t n_tiles, int C) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (e >= LS) return;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int64_t t = c;
  for (; t + 3 * C < n_tiles; t += 4 * C) {
    s0 += part[t * LS + e];
    s1 += part[(t + C) * LS + e];
    s2 += part[(t + 2 * C) * LS + e];
    s3 += part[(t + 3 * C) * LS + e];
  }
  for (; t < n_tiles; t += C) s0 += part[t * LS + e];
  part[(n_tiles + c) * LS + e] = (s0 + s1) + (s2 + s3);
}

__global__ void __launch_bounds__(256) rnvp_wgrad_stage2(RnvpDev f, int l, const float* __restrict__ part,
                                                         int64_t n_tiles, int C, float* __restrict__ grad) {
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
  for (int c = 0; c < C; ++c) s += part[(n_tiles + c) * LS + e];
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
  if (partial_per_tile) *partial_per_tile = rn_layer_size(f.d, f.h);  // rows: tiles + chunk rows
  if (tile) *tile = RN_TS;
  return FMC_OK;
}

extern "C" int fmc_rnvp_pack(void* stream, const fmc_rnvp* r) {
  RnvpDev f;
  int rc = check_rnvp(r, f, true);
  if (rc) return rc;
  const int64_t n = rn_pk_mlp(f.d, f.h);
  const unsigned gx = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
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

extern "C" int fmc_rnvp_weight_grad(void* stream, const fmc_rnvp* r, int32_t layer, float* partials,
                                    int64_t n_tiles, int64_t partial_rows, float* grad) {
  RnvpDev f;
  int rc = check_rnvp(r, f, false);
  if (rc) return rc;
  FMC_ARG(layer >= 0 && layer < f.L, "layer %d out of range", layer);
  FMC_ARG(partials && grad, "NULL pointer");
  FMC_ARG(n_tiles >= 1, "n_tiles must be >= 1");
  const int C = rn_chunks(n_tiles);
  FMC_ARG(partial_rows >= n_tiles + C, "partials hold %lld rows, %lld needed (tiles + %d reduction rows)",
          (long long)partial_rows, (long long)(n_tiles + C), C);
  const int64_t LS = rn_layer_size(f.d, f.h);
  const unsigned gx = (unsigned)((LS + 255) / 256);
  hipLaunchKernelGGL(rnvp_wgrad_stage1, dim3(gx, C), dim3(256), 0, (hipStream_t)stream, LS, partials, n_tiles, C);
  if ((rc = check_launch("realnvp weight gradient (tile sums)"))) return rc;
  hipLaunchKernelGGL(rnvp_wgrad_stage2, dim3(gx), dim3(256), 0, (hipStream_t)stream, f, layer, partials, n_tiles,
                     C, grad);
  return check_launch("realnvp weight gradient");
}
