// MaskedCouplingRQSpline flow on gfx950: packed MFMA conditioner GEMMs with
// the rational-quadratic spline fused as the epilogue of the last GEMM.
//
//   flow_pass_kernel      all layers of a forward / inverse pass for a tile of
//                         32*MB samples (activations stay in LDS)
//   flow_bwd_kernel       adjoint of one forward layer (recomputes conditioner
//                         and spline bin parameters, hand-derived spline VJP),
//                         stores activations / gradients (fragment-blocked)
//   flow_wgrad_kernel     weight, bias and ScalarAffine gradients: split batch
//                         reduction on MFMA
//   flow_pack_kernel      master FP32 parameters -> MFMA operand images
//   adamw_* kernels       optax clip_by_global_norm(1) + adamw
//
// Reference: src/flowMC/resource/model/nf_model/rqSpline.py (spline :20-239,
// RQSpline :242-361, MaskedCouplingRQSpline :364-504), common.py (MLP
// :68-124, MaskedCouplingLayer :127-168, ScalarAffine :211-240, Gaussian
// :243-293), nf_model/base.py:98-125 (loss / train_step),
// src/flowMC/resource/optimizer.py:19-22.
#include "fmc_common.h"
#include "fmc_flow_layout.h"

namespace fmc {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// --------------------------------------------------------- MFMA traits ---
// 32x32 output block per wavefront. A operand from LDS (row-major, stride
// lda), B operand from a packed "B-fragment image" in global memory: block
// (nb, kb) holds 64 lanes x EPL contiguous elements, lane l carrying column
// nb*32 + l%32 and k = kb*KSTEP + (l/32)*EPL + e.
// Output element v of lane l is row 8*(v/4) + 4*(l/32) + v%4, column l%32.
struct PF32 {
  using T = float;
  static constexpr int EPL = 1, KSTEP = 2, PAD = 1;
  static __device__ __forceinline__ float load_a(const T* As, int lda, int row0, int kb, int lane) {
    return As[(row0 + (lane & 31)) * lda + kb * 2 + (lane >> 5)];
  }
  static __device__ __forceinline__ float load_b(const T* Bp, int KB, int nb, int kb, int lane) {
    return Bp[(int64_t)(nb * KB + kb) * 64 + lane];
  }
  static __device__ __forceinline__ f32x16_t mma(float a, float b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  // fragment of row block rb, batch slice s of the fragment-blocked store (ns slices per row block)
  static __device__ __forceinline__ float load_g(const T* base, int64_t ns, int rb, int64_t s, int lane) {
    return base[((int64_t)rb * ns + s) * 64 + lane];
  }
  static __device__ __forceinline__ T cvt(float x) { return x; }
  static __device__ __forceinline__ float f(T x) { return x; }
  static __device__ __forceinline__ float hsum(float x) { return x; }
  static __device__ __forceinline__ float act(float x) { return tanhf(x); }
  // 4 consecutive elements (rows are not 16-byte aligned: 4 scalar stores)
  static __device__ __forceinline__ void store4(T* dst, float a, float b, float c, float d) {
    dst[0] = a;
    dst[1] = b;
    dst[2] = c;
    dst[3] = d;
  }
};

struct PBF16 {
  using T = __bf16;
  static constexpr int EPL = 8, KSTEP = 16, PAD = 8;
  static __device__ __forceinline__ bf16x8_t load_a(const T* As, int lda, int row0, int kb, int lane) {
    return *reinterpret_cast<const bf16x8_t*>(As + (row0 + (lane & 31)) * lda + kb * 16 + (lane >> 5) * 8);
  }
  static __device__ __forceinline__ bf16x8_t load_b(const T* Bp, int KB, int nb, int kb, int lane) {
    return *reinterpret_cast<const bf16x8_t*>(Bp + ((int64_t)(nb * KB + kb) * 64 + lane) * 8);
  }
  static __device__ __forceinline__ f32x16_t mma(bf16x8_t a, bf16x8_t b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ bf16x8_t load_g(const T* base, int64_t ns, int rb, int64_t s, int lane) {
    return *reinterpret_cast<const bf16x8_t*>(base + (((int64_t)rb * ns + s) * 64 + lane) * 8);
  }
  static __device__ __forceinline__ T cvt(float x) { return (__bf16)x; }
  static __device__ __forceinline__ float f(T x) { return (float)x; }
  // 4 consecutive elements, one 8-byte store (dst is 8-byte aligned)
  static __device__ __forceinline__ void store4(T* dst, float a, float b, float c, float d) {
    bf16x4_t v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    v[2] = (__bf16)c;
    v[3] = (__bf16)d;
    *reinterpret_cast<bf16x4_t*>(dst) = v;
  }
  static __device__ __forceinline__ float hsum(bf16x8_t x) {
    return (((float)x[0] + (float)x[1]) + ((float)x[2] + (float)x[3])) +
           (((float)x[4] + (float)x[5]) + ((float)x[6] + (float)x[7]));
  }
  static __device__ __forceinline__ float act(float x) {
    // 1 - 2 / (e^{2x} + 1); the reciprocal from v_rcp_f32 refined by one
    // Newton step (correctly rounded but for rare last-bit cases)
    const float den = __expf(2.f * x) + 1.f;
    float r = __builtin_amdgcn_rcpf(den);
    r = fmaf(fmaf(-den, r, 1.f), r, r);
    return fmaf(-2.f, r, 1.f);
  }
};

__device__ __forceinline__ int out_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// Epilogue of a transposed hidden-layer GEMM (output rows = features, columns
// = samples): lane l holds, for sample r*32 + l%32, features
// fb*32 + 8g + 4(l/32) + i (g, i = 0..3) in acc[4g + i]; writes
// tanh(acc + bias) as 4-element runs into the [sample][feature] LDS rows.
template <class PT, int MB>
__device__ __forceinline__ void hidden_epilogue_tr(typename PT::T* hb, int lh, const float* bias, int fb,
                                                   const f32x16_t (&acc)[MB], int lane) {
  const int f0 = fb * 32 + 4 * (lane >> 5);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 b4 = *reinterpret_cast<const float4*>(bias + f0 + 8 * g);
#pragma unroll
    for (int r = 0; r < MB; ++r)
      PT::store4(hb + (r * 32 + (lane & 31)) * lh + f0 + 8 * g, PT::act(acc[r][4 * g] + b4.x),
                 PT::act(acc[r][4 * g + 1] + b4.y), PT::act(acc[r][4 * g + 2] + b4.z),
                 PT::act(acc[r][4 * g + 3] + b4.w));
  }
}

// a / b from v_rcp_f32 with one Newton step on the reciprocal and one on the
// quotient: the IEEE-rounded quotient but for rare last-bit cases, at half
// the instructions of the full division sequence (operands here are finite
// and normal; b = 0 gives a NaN that the callers clamp, as the IEEE
// division's 0/0 would)
__device__ __forceinline__ float ndiv(float a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  r = fmaf(fmaf(-b, r, 1.f), r, r);
  const float q = a * r;
  return fmaf(fmaf(-b, q, a), r, q);
}

// (m, j) walk over an MT x n tile by the 256 threads of a workgroup, without
// a division per element: start at tid, advance by 256 elements.
struct TileWalk {
  int m, j, dm, dj, n;
  __device__ TileWalk(int tid, int n_) : n(n_) {
    m = tid / n;
    j = tid - m * n;
    dm = 256 / n;
    dj = 256 - dm * n;
  }
  __device__ void next() {
    j += dj;
    m += dm;
    if (j >= n) {
      j -= n;
      ++m;
    }
  }
};

// One 32x32 block: acc += A[row0:row0+32, kb0*KSTEP : kb1*KSTEP] * B-image block column nb.
template <class PT>
__device__ __forceinline__ void mfma_block(const typename PT::T* As, int lda, int row0, const typename PT::T* Bp,
                                           int KBtot, int nb, int kb0, int kb1, int a_kb0, f32x16_t& acc, int lane) {
#pragma unroll 4
  for (int kb = kb0; kb < kb1; ++kb) {
    auto a = PT::load_a(As, lda, row0, kb - kb0 + a_kb0, lane);
    auto b = PT::load_b(Bp, KBtot, nb, kb, lane);
    acc = PT::mma(a, b, acc);
  }
}

// Panel of MB row blocks x up to CB column blocks (nb0, nb0 + nbstep, ...):
// every A fragment feeds CB MFMAs, every B fragment MB MFMAs, and the B
// fragments (global / L2) are prefetched P k-steps ahead in a register ring
// so their latency hides behind the MFMAs of the previous steps.
// TR = true swaps the operands: the packed image is the A operand (output
// rows = its rows), the LDS rows are the B operand (output columns).
template <class PT, int MB, int CB, int P, bool TR = false>
__device__ __forceinline__ void mfma_panel(const typename PT::T* As, int lda, const typename PT::T* Bp, int KBtot,
                                           int nb0, int nbstep, int ncb, int kb0, int kb1, int a_kb0,
                                           f32x16_t (&acc)[CB][MB], int lane) {
  using BF = decltype(PT::load_b(Bp, 0, 0, 0, 0));
  using AF = decltype(PT::load_a(As, 0, 0, 0, 0));
  BF ring[P][CB];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int j = 0; j < CB; ++j)
      if (kb0 + p < kb1 && j < ncb) ring[p][j] = PT::load_b(Bp, KBtot, nb0 + j * nbstep, kb0 + p, lane);
  for (int kb = kb0; kb < kb1; kb += P) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int k = kb + p;
      if (k < kb1) {
        AF a[MB];
#pragma unroll
        for (int r = 0; r < MB; ++r) a[r] = PT::load_a(As, lda, r * 32, k - kb0 + a_kb0, lane);
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          if (j < ncb) {
#pragma unroll
            for (int r = 0; r < MB; ++r)
              acc[j][r] = TR ? PT::mma(ring[p][j], a[r], acc[j][r]) : PT::mma(a[r], ring[p][j], acc[j][r]);
            if (k + P < kb1) ring[p][j] = PT::load_b(Bp, KBtot, nb0 + j * nbstep, k + P, lane);
          }
        }
      }
    }
  }
}

constexpr int PF_DEPTH = 4;  // B-fragment prefetch depth (k-steps); the kernels run 2-3 waves per SIMD
// the pass kernel fits an 8-deep ring in its three-waves-per-SIMD budget (167 VGPRs)
constexpr int PASS_PF = 8;

// --------------------------------------------------------------- flow ----
struct FlowDev {
  FlowDims dm;
  const float* theta;
  const void* packed;
  const float* pbias;
  float lo, hi, mb, ms, off;  // spline range, min bin size, min knot slope, softplus offset
};

__device__ __forceinline__ float softplus(float v) { return fmaxf(v, 0.f) + log1pf(expf(-fabsf(v))); }
__device__ __forceinline__ float sigmoid(float v) { return 1.f / (1.f + expf(-v)); }

// Rational-quadratic spline of one scalar given its 3K+1 unnormalised
// parameters pp (widths | heights | slopes): rqSpline.py:20-128.
__device__ void spline_fwd(const float* pp, const FlowDev& f, float x, float& y, float& ld) {
  const int K = f.dm.K;
  if (x <= f.lo) {
    float d0 = softplus(pp[2 * K] + f.off) + f.ms;
    y = (x - f.lo) * d0 + f.lo;
    ld = logf(d0);
    return;
  }
  if (x >= f.hi) {
    float dK = softplus(pp[3 * K] + f.off) + f.ms;
    y = (x - f.hi) * dK + f.hi;
    ld = logf(dK);
    return;
  }
  float mw = -INFINITY, mh = -INFINITY;
  for (int i = 0; i < K; ++i) {
    mw = fmaxf(mw, pp[i]);
    mh = fmaxf(mh, pp[K + i]);
  }
  float sw = 0.f, sh = 0.f;
  for (int i = 0; i < K; ++i) {
    sw += expf(pp[i] - mw);
    sh += expf(pp[K + i] - mh);
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float cw = span / sw, ch = span / sh;
  float xk = f.lo, yk = f.lo, xk1 = f.hi, yk1 = f.hi;
  int k = 0;
  for (int i = 0; i < K; ++i) {
    float xn = (i == K - 1) ? f.hi : xk + (expf(pp[i] - mw) * cw + f.mb);
    float yn = (i == K - 1) ? f.hi : yk + (expf(pp[K + i] - mh) * ch + f.mb);
    if (x < xn) {
      k = i;
      xk1 = xn;
      yk1 = yn;
      break;
    }
    xk = xn;
    yk = yn;
  }
  float dk = softplus(pp[2 * K + k] + f.off) + f.ms;
  float dk1 = softplus(pp[2 * K + k + 1] + f.off) + f.ms;
  float w = xk1 - xk, hh = yk1 - yk, s = hh / w;
  float z = fminf(fmaxf((x - xk) / w, 0.f), 1.f);
  float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  float st = dk1 + dk - 2.f * s;
  float num = hh * (s * z2 + dk * zz);
  float den = s + st * zz;
  y = yk + num / den;
  ld = 2.f * logf(s) + logf(dk1 * z2 + 2.f * s * zz + dk * q) - 2.f * logf(den);
}

// Inverse spline: rqSpline.py:131-239 (stable quadratic root).
__device__ void spline_inv(const float* pp, const FlowDev& f, float yv, float& x, float& ld) {
  const int K = f.dm.K;
  if (yv <= f.lo) {
    float d0 = softplus(pp[2 * K] + f.off) + f.ms;
    x = (yv - f.lo) / d0 + f.lo;
    ld = -logf(d0);
    return;
  }
  if (yv >= f.hi) {
    float dK = softplus(pp[3 * K] + f.off) + f.ms;
    x = (yv - f.hi) / dK + f.hi;
    ld = -logf(dK);
    return;
  }
  float mw = -INFINITY, mh = -INFINITY;
  for (int i = 0; i < K; ++i) {
    mw = fmaxf(mw, pp[i]);
    mh = fmaxf(mh, pp[K + i]);
  }
  float sw = 0.f, sh = 0.f;
  for (int i = 0; i < K; ++i) {
    sw += expf(pp[i] - mw);
    sh += expf(pp[K + i] - mh);
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float cw = span / sw, ch = span / sh;
  float xk = f.lo, yk = f.lo, xk1 = f.hi, yk1 = f.hi;
  int k = 0;
  for (int i = 0; i < K; ++i) {
    float xn = (i == K - 1) ? f.hi : xk + (expf(pp[i] - mw) * cw + f.mb);
    float yn = (i == K - 1) ? f.hi : yk + (expf(pp[K + i] - mh) * ch + f.mb);
    if (yv < yn) {
      k = i;
      xk1 = xn;
      yk1 = yn;
      break;
    }
    xk = xn;
    yk = yn;
  }
  float dk = softplus(pp[2 * K + k] + f.off) + f.ms;
  float dk1 = softplus(pp[2 * K + k + 1] + f.off) + f.ms;
  float w = xk1 - xk, hh = yk1 - yk, s = hh / w;
  float wv = fminf(fmaxf((yv - yk) / hh, 0.f), 1.f);
  float st = dk1 + dk - 2.f * s;
  float c = -s * wv;
  float b = dk - st * wv;
  float a = s - b;
  float disc = b * b - 4.f * a * c;
  float sq = disc > 0.f ? sqrtf(fmaxf(disc, 1.17549435e-38f)) : 0.f;
  float z = (b >= 0.f) ? (2.f * c) / (-b - sq) : (-b + sq) / (2.f * a);
  z = fminf(fmaxf(z, 0.f), 1.f);
  x = w * z + xk;
  float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  float den = s + st * zz;
  ld = -2.f * logf(s) - logf(dk1 * z2 + 2.f * s * zz + dk * q) + 2.f * logf(den);
}

// Forward spline + vector-Jacobian product. Given upstream gy (dL/dy) and gl
// (dL/dlogdet) writes dL/d(unnormalised params) to go[0..3K] and returns dL/dx.
template <class T, class CV>
__device__ float spline_vjp(const float* pp, const FlowDev& f, float x, float gy, float gl, T* go, CV cvt) {
  const int K = f.dm.K;
  for (int i = 0; i < 3 * K + 1; ++i) go[i] = cvt(0.f);
  if (x <= f.lo || x >= f.hi) {
    const int idx = (x <= f.lo) ? 2 * K : 3 * K;
    const float edge = (x <= f.lo) ? f.lo : f.hi;
    float u = pp[idx] + f.off;
    float dd = softplus(u) + f.ms;
    float gd = gy * (x - edge) + gl / dd;
    go[idx] = cvt(gd * sigmoid(u));
    return gy * dd;
  }
  float mw = -INFINITY, mh = -INFINITY;
  for (int i = 0; i < K; ++i) {
    mw = fmaxf(mw, pp[i]);
    mh = fmaxf(mh, pp[K + i]);
  }
  float sw = 0.f, sh = 0.f;
  for (int i = 0; i < K; ++i) {
    sw += expf(pp[i] - mw);
    sh += expf(pp[K + i] - mh);
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float cw = span / sw, ch = span / sh;
  float xk = f.lo, yk = f.lo, xk1 = f.hi, yk1 = f.hi;
  float pw_lt = 0.f, ph_lt = 0.f;  // sum of softmax probs of bins < k
  float pwk = 0.f, phk = 0.f;
  int k = 0;
  for (int i = 0; i < K; ++i) {
    float ew = expf(pp[i] - mw), eh = expf(pp[K + i] - mh);
    float xn = (i == K - 1) ? f.hi : xk + (ew * cw + f.mb);
    float yn = (i == K - 1) ? f.hi : yk + (eh * ch + f.mb);
    if (x < xn) {
      k = i;
      xk1 = xn;
      yk1 = yn;
      pwk = ew / sw;
      phk = eh / sh;
      break;
    }
    pw_lt += ew / sw;
    ph_lt += eh / sh;
    xk = xn;
    yk = yn;
  }
  const float uk = pp[2 * K + k] + f.off, uk1 = pp[2 * K + k + 1] + f.off;
  const float dk = softplus(uk) + f.ms, dk1 = softplus(uk1) + f.ms;
  const float w = xk1 - xk, hh = yk1 - yk, s = hh / w;
  const float t = (x - xk) / w;
  const float z = fminf(fmaxf(t, 0.f), 1.f);
  const float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  const float st = dk1 + dk - 2.f * s;
  const float num = hh * (s * z2 + dk * zz);
  const float den = s + st * zz;
  const float A = dk1 * z2 + 2.f * s * zz + dk * q;
  // reverse sweep of y = yk + num/den, ld = 2 log s + log A - 2 log den
  float g_num = gy / den;
  float g_den = -gy * num / (den * den) - 2.f * gl / den;
  float g_s = 2.f * gl / s;
  float g_A = gl / A;
  float g_yk = gy;
  float g_dk1 = g_A * z2, g_z2 = g_A * dk1, g_zz = g_A * 2.f * s, g_dk = g_A * q, g_q = g_A * dk;
  g_s += g_A * 2.f * zz;
  g_s += g_den;
  float g_st = g_den * zz;
  g_zz += g_den * st;
  float g_hh = g_num * (s * z2 + dk * zz);
  g_s += g_num * hh * z2;
  g_z2 += g_num * hh * s;
  g_dk += g_num * hh * zz;
  g_zz += g_num * hh * dk;
  g_dk1 += g_st;
  g_dk += g_st;
  g_s += -2.f * g_st;
  g_z2 -= g_zz;
  float g_z = g_zz + g_z2 * 2.f * z - g_q * 2.f * (1.f - z);
  float g_t = (t > 0.f && t < 1.f) ? g_z : 0.f;
  float g_x = g_t / w;
  float g_xk = -g_t / w;
  float g_w = -g_t * t / w;
  g_hh += g_s / w;
  g_w += -g_s * s / w;
  float g_yk1 = g_hh;
  g_yk -= g_hh;
  float g_xk1 = g_w;
  g_xk -= g_w;
  // knots -> softmax logits: x_i = lo + sum_{m<i} w_m for 1 <= i <= K-1
  const float G1x = (k >= 1) ? g_xk : 0.f, G2x = (k + 1 <= K - 1) ? g_xk1 : 0.f;
  const float G1y = (k >= 1) ? g_yk : 0.f, G2y = (k + 1 <= K - 1) ? g_yk1 : 0.f;
  const float dotw = span * (G1x * pw_lt + G2x * (pw_lt + pwk));
  const float doth = span * (G1y * ph_lt + G2y * (ph_lt + phk));
  for (int m = 0; m < K; ++m) {
    float gwm = span * ((m < k ? G1x : 0.f) + (m <= k ? G2x : 0.f));
    float ghm = span * ((m < k ? G1y : 0.f) + (m <= k ? G2y : 0.f));
    float pw = expf(pp[m] - mw) / sw, ph = expf(pp[K + m] - mh) / sh;
    go[m] = cvt(pw * (gwm - dotw));
    go[K + m] = cvt(ph * (ghm - doth));
  }
  go[2 * K + k] = cvt(g_dk * sigmoid(uk));
  go[2 * K + k + 1] = cvt(g_dk1 * sigmoid(uk1));
  return g_x;
}

// ------------------------------------------------ register spline -----
// The same spline with its 3K+1 parameters in registers, in the slot order of
// the register-spline W3 layout (widths at 0..9, heights at 10..19, slopes of
// knots 0..K-1 at 20..29, slope of knot K at 30; K <= 10). Every loop is
// unrolled over the 10 slots and predicated on K, so all register indices are
// static. The bin search keeps the first bin whose right knot exceeds the
// argument, as spline_fwd / spline_inv do.
struct RsBin {
  float xk, xk1, yk, yk1, uk, uk1;
};

__device__ __forceinline__ float rs_softplus(float v) { return fmaxf(v, 0.f) + __logf(1.f + __expf(-fabsf(v))); }

// knots of the bin holding t (t along x when INV is false, along y otherwise)
// KT > 0: the bin count is the compile-time constant KT (loops unrolled to
// exactly KT, no predication); KT == 0: runtime K <= 10, loops predicated.
template <bool INV, int KT>
__device__ __forceinline__ RsBin rs_bin(const float (&pp)[32], int Kr, const FlowDev& f, float t) {
  constexpr int N = KT > 0 ? KT : 10;
  const int K = KT > 0 ? KT : Kr;
  float mw = -INFINITY, mh = -INFINITY;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (KT > 0 || i < K) {
      mw = fmaxf(mw, pp[i]);
      mh = fmaxf(mh, pp[10 + i]);
    }
  float ew[10], eh[10], sw = 0.f, sh = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ew[i] = (KT > 0 || i < K) ? __expf(pp[i] - mw) : 0.f;
    eh[i] = (KT > 0 || i < K) ? __expf(pp[10 + i] - mh) : 0.f;
    sw += ew[i];
    sh += eh[i];
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float cw = ndiv(span, sw), ch = ndiv(span, sh);
  RsBin b{f.lo, f.hi, f.lo, f.hi, pp[20], pp[21]};
  float xc = f.lo, yc = f.lo;
  bool found = false;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (KT > 0 || i < K) {
      const float xn = (i == K - 1) ? f.hi : xc + (ew[i] * cw + f.mb);
      const float yn = (i == K - 1) ? f.hi : yc + (eh[i] * ch + f.mb);
      const bool hit = !found && (INV ? (t < yn) : (t < xn));
      if (hit) {
        b.xk = xc;
        b.yk = yc;
        b.xk1 = xn;
        b.yk1 = yn;
        b.uk = pp[20 + i];
        b.uk1 = (i == K - 1) ? pp[30] : pp[21 + i];
      }
      found = found || hit;
      xc = xn;
      yc = yn;
    }
  }
  return b;
}

template <int KT>
__device__ __forceinline__ void rs_fwd(const float (&pp)[32], int K, const FlowDev& f, float x, float& y, float& ld) {
  if (x <= f.lo || x >= f.hi) {
    const float u = (x <= f.lo) ? pp[20] : pp[30];
    const float e = (x <= f.lo) ? f.lo : f.hi;
    const float dd = rs_softplus(u + f.off) + f.ms;
    y = (x - e) * dd + e;
    ld = __logf(dd);
    return;
  }
  const RsBin b = rs_bin<false, KT>(pp, K, f, x);
  const float dk = rs_softplus(b.uk + f.off) + f.ms, dk1 = rs_softplus(b.uk1 + f.off) + f.ms;
  const float w = b.xk1 - b.xk, hh = b.yk1 - b.yk, s = ndiv(hh, w);
  const float z = fminf(fmaxf(ndiv(x - b.xk, w), 0.f), 1.f);
  const float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  const float st = dk1 + dk - 2.f * s;
  const float num = hh * (s * z2 + dk * zz);
  const float den = s + st * zz;
  y = b.yk + ndiv(num, den);
  ld = 2.f * __logf(s) + __logf(dk1 * z2 + 2.f * s * zz + dk * q) - 2.f * __logf(den);
}

template <int KT>
__device__ __forceinline__ void rs_inv(const float (&pp)[32], int K, const FlowDev& f, float yv, float& x, float& ld) {
  if (yv <= f.lo || yv >= f.hi) {
    const float u = (yv <= f.lo) ? pp[20] : pp[30];
    const float e = (yv <= f.lo) ? f.lo : f.hi;
    const float dd = rs_softplus(u + f.off) + f.ms;
    x = (yv - e) / dd + e;
    ld = -__logf(dd);
    return;
  }
  const RsBin b = rs_bin<true, KT>(pp, K, f, yv);
  const float dk = rs_softplus(b.uk + f.off) + f.ms, dk1 = rs_softplus(b.uk1 + f.off) + f.ms;
  const float w = b.xk1 - b.xk, hh = b.yk1 - b.yk, s = ndiv(hh, w);
  const float wv = fminf(fmaxf(ndiv(yv - b.yk, hh), 0.f), 1.f);
  const float st = dk1 + dk - 2.f * s;
  const float c = -s * wv;
  const float bq = dk - st * wv;
  const float a = s - bq;
  const float disc = bq * bq - 4.f * a * c;
  const float sq = disc > 0.f ? sqrtf(fmaxf(disc, 1.17549435e-38f)) : 0.f;
  float z = (bq >= 0.f) ? ndiv(2.f * c, -bq - sq) : ndiv(-bq + sq, 2.f * a);
  z = fminf(fmaxf(z, 0.f), 1.f);
  x = w * z + b.xk;
  const float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  const float den = s + st * zz;
  ld = -2.f * __logf(s) - __logf(dk1 * z2 + 2.f * s * zz + dk * q) + 2.f * __logf(den);
}

// Register-spline forward + vector-Jacobian product (the reverse sweep of
// spline_vjp): given dL/dy (gy) and dL/dlogdet (gl), writes dL/d(slot
// logits) to go (zero in padding slots) and returns dL/dx.
template <int KT>
__device__ __forceinline__ float rs_vjp(const float (&pp)[32], int Kr, const FlowDev& f, float x, float gy, float gl,
                                        float (&go)[32]) {
  constexpr int N = KT > 0 ? KT : 10;
  const int K = KT > 0 ? KT : Kr;
#pragma unroll
  for (int i = 0; i < 32; ++i) go[i] = 0.f;
  if (x <= f.lo || x >= f.hi) {
    const bool left = x <= f.lo;
    const float u = (left ? pp[20] : pp[30]) + f.off;
    const float dd = rs_softplus(u) + f.ms;
    const float gd = gy * (x - (left ? f.lo : f.hi)) + ndiv(gl, dd);
    const float gu = ndiv(gd, 1.f + __expf(-u));
    go[20] = left ? gu : 0.f;
    go[30] = left ? 0.f : gu;
    return gy * dd;
  }
  float mw = -INFINITY, mh = -INFINITY;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (KT > 0 || i < K) {
      mw = fmaxf(mw, pp[i]);
      mh = fmaxf(mh, pp[10 + i]);
    }
  float ew[10], eh[10], sw = 0.f, sh = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ew[i] = (KT > 0 || i < K) ? __expf(pp[i] - mw) : 0.f;
    eh[i] = (KT > 0 || i < K) ? __expf(pp[10 + i] - mh) : 0.f;
    sw += ew[i];
    sh += eh[i];
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float isw = ndiv(1.f, sw), ish = ndiv(1.f, sh);
  const float cw = span * isw, ch = span * ish;
  float xk = f.lo, yk = f.lo, xk1 = f.hi, yk1 = f.hi, uk = pp[20], uk1 = pp[21];
  float pw_lt = 0.f, ph_lt = 0.f, pwk = 0.f, phk = 0.f;  // softmax mass of bins < k, and of bin k
  float xc = f.lo, yc = f.lo;
  int k = 0;
  bool found = false;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (KT > 0 || i < K) {
      const float xn = (i == K - 1) ? f.hi : xc + (ew[i] * cw + f.mb);
      const float yn = (i == K - 1) ? f.hi : yc + (eh[i] * ch + f.mb);
      const bool hit = !found && (x < xn);
      if (hit) {
        k = i;
        xk = xc;
        yk = yc;
        xk1 = xn;
        yk1 = yn;
        uk = pp[20 + i];
        uk1 = (i == K - 1) ? pp[30] : pp[21 + i];
        pwk = ew[i] * isw;
        phk = eh[i] * ish;
      }
      if (!found && !hit) {
        pw_lt += ew[i] * isw;
        ph_lt += eh[i] * ish;
      }
      found = found || hit;
      xc = xn;
      yc = yn;
    }
  }
  uk += f.off;
  uk1 += f.off;
  const float dk = rs_softplus(uk) + f.ms, dk1 = rs_softplus(uk1) + f.ms;
  const float w = xk1 - xk, hh = yk1 - yk, s = ndiv(hh, w);
  const float t = ndiv(x - xk, w);
  const float z = fminf(fmaxf(t, 0.f), 1.f);
  const float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  const float st = dk1 + dk - 2.f * s;
  const float num = hh * (s * z2 + dk * zz);
  const float den = s + st * zz;
  const float A = dk1 * z2 + 2.f * s * zz + dk * q;
  // reverse sweep of y = yk + num/den, ld = 2 log s + log A - 2 log den
  const float g_num = ndiv(gy, den);
  const float g_den = -ndiv(gy * num, den * den) - ndiv(2.f * gl, den);
  float g_s = ndiv(2.f * gl, s);
  const float g_A = ndiv(gl, A);
  float g_yk = gy;
  float g_dk1 = g_A * z2, g_z2 = g_A * dk1, g_zz = g_A * 2.f * s, g_dk = g_A * q;
  const float g_q = g_A * dk;
  g_s += g_A * 2.f * zz;
  g_s += g_den;
  const float g_st = g_den * zz;
  g_zz += g_den * st;
  float g_hh = g_num * (s * z2 + dk * zz);
  g_s += g_num * hh * z2;
  g_z2 += g_num * hh * s;
  g_dk += g_num * hh * zz;
  g_zz += g_num * hh * dk;
  g_dk1 += g_st;
  g_dk += g_st;
  g_s += -2.f * g_st;
  g_z2 -= g_zz;
  const float g_z = g_zz + g_z2 * 2.f * z - g_q * 2.f * (1.f - z);
  const float g_t = (t > 0.f && t < 1.f) ? g_z : 0.f;
  const float g_x = ndiv(g_t, w);
  float g_xk = -g_x;
  float g_w = -ndiv(g_t * t, w);
  g_hh += ndiv(g_s, w);
  g_w += -ndiv(g_s * s, w);
  const float g_yk1 = g_hh;
  g_yk -= g_hh;
  const float g_xk1 = g_w;
  g_xk -= g_w;
  // knots -> softmax logits: x_i = lo + sum_{m<i} w_m for 1 <= i <= K-1
  const float G1x = (k >= 1) ? g_xk : 0.f, G2x = (k + 1 <= K - 1) ? g_xk1 : 0.f;
  const float G1y = (k >= 1) ? g_yk : 0.f, G2y = (k + 1 <= K - 1) ? g_yk1 : 0.f;
  const float dotw = span * (G1x * pw_lt + G2x * (pw_lt + pwk));
  const float doth = span * (G1y * ph_lt + G2y * (ph_lt + phk));
#pragma unroll
  for (int m = 0; m < N; ++m) {
    if (KT > 0 || m < K) {
      const float gwm = span * ((m < k ? G1x : 0.f) + (m <= k ? G2x : 0.f));
      const float ghm = span * ((m < k ? G1y : 0.f) + (m <= k ? G2y : 0.f));
      go[m] = ew[m] * isw * (gwm - dotw);
      go[10 + m] = eh[m] * ish * (ghm - doth);
    }
  }
  const float gk = ndiv(g_dk, 1.f + __expf(-uk)), gk1 = ndiv(g_dk1, 1.f + __expf(-uk1));
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (KT > 0 || i < K) go[20 + i] = (i == k ? gk : 0.f) + (i == k + 1 ? gk1 : 0.f);
  go[30] = (k + 1 == K) ? gk1 : 0.f;
  return g_x;
}

// ------------------------------------------------------- pass kernel -----
struct PassArgs {
  int dir;    // 0 forward, 1 inverse (reference layer order), 2 exact inverse
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

enum { FL_STD = 1, FL_BASE = 2, FL_SAMPLE = 4 };

template <class PT>
struct PassSmem {
  // byte offsets inside dynamic LDS
  int xs, ldr, ldp, r1, r2, total;
  int ldx, lh1, lh2, lo_, la0;
  __host__ __device__ PassSmem(const FlowDims& dm, int MT) {
    ldx = dm.d + 1;
    lh1 = dm.h1p() + PT::PAD;
    lh2 = dm.h2p() + PT::PAD;
    lo_ = dm.ncf() + 1;
    la0 = rup(dm.kc(1) > dm.kc(0) ? dm.kc(1) : dm.kc(0), PT::KSTEP) + PT::PAD;
    int o = 0;
    xs = o;
    o += rup(MT * ldx * 4, 16);
    ldr = o;
    o += rup(MT * 4, 16);
    ldp = o;  // per-dim log-dets of a chunk, or the 4 waves' partial log-dets
    o += rup(MT * (dm.G() > 4 ? dm.G() : 4) * 4, 16);
    r1 = o;
    int s1 = MT * lh1 * (int)sizeof(typename PT::T), s1b = MT * lo_ * 4;
    o += rup(s1 > s1b ? s1 : s1b, 16);
    r2 = o;
    int s2 = MT * lh2 * (int)sizeof(typename PT::T), s2b = MT * la0 * (int)sizeof(typename PT::T);
    o += rup(s2 > s2b ? s2 : s2b, 16);
    total = o;
  }
};

// 32-sample tiles (MB = 1) with the register budget held to three waves per
// SIMD: three workgroups share a CU (LDS ~43 KB each at d=64, [256,256]) and
// hide each other's LDS / L2 latency; measured faster than 64-sample tiles
// with more B-fragment reuse per wave.
template <class PT, int MB, int KT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
flow_pass_kernel(FlowDev f, PassArgs a) {
  using T = typename PT::T;
  constexpr int MT = 32 * MB;
  extern __shared__ __align__(16) unsigned char smem[];
  const FlowDims& dm = f.dm;
  PassSmem<PT> sm(dm, MT);
  float* xs = reinterpret_cast<float*>(smem + sm.xs);
  float* ldr = reinterpret_cast<float*>(smem + sm.ldr);
  float* ldp = reinterpret_cast<float*>(smem + sm.ldp);
  T* hb1 = reinterpret_cast<T*>(smem + sm.r1);
  float* obuf = reinterpret_cast<float*>(smem + sm.r1);
  T* hb2 = reinterpret_cast<T*>(smem + sm.r2);
  T* a0 = reinterpret_cast<T*>(smem + sm.r2);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int d = dm.d, ldx = sm.ldx;
  const int64_t m0 = (int64_t)blockIdx.x * MT;
  const float* th = f.theta;
  const T* pk = reinterpret_cast<const T*>(f.packed);

  // ---- load tile (+ standardisation / base sampling)
  for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
    const int m = tw.m, j = tw.j;
    int64_t row = m0 + m;
    float v = 0.f;
    if (row < a.n) {
      if ((a.flags & FL_SAMPLE) && a.x_in == nullptr) {
        float4 z = philox_normal4(a.key0, a.key1, (uint32_t)(j >> 2), (uint32_t)(a.sample0 + row), 0u, RNG_FLOW_BASE);
        const int e = j & 3;
        v = (e == 0) ? z.x : (e == 1) ? z.y : (e == 2) ? z.z : z.w;
      } else {
        v = a.x_in[row * d + j];
      }
      if (a.dir == 0 && (a.flags & FL_STD)) {
        v = (v - th[dm.th_data_mean() + j]) / sqrtf(th[dm.th_data_cov() + (int64_t)j * d + j]);
      } else if (a.dir != 0 && (a.flags & FL_SAMPLE)) {
        v = th[dm.th_base_mean() + j] + sqrtf(th[dm.th_base_cov() + (int64_t)j * d + j]) * v;
      }
    }
    xs[m * ldx + j] = v;
  }
  __syncthreads();
  // exact inverse with BASE_LOGP: the pass returns the flow log-density of
  // its output, log N(z) - sum of the inverse log-dets; start from -log N(z)
  const bool inv_logq = (a.dir == 2) && (a.flags & FL_BASE);
  for (int m = tid; m < MT; m += 256) {
    float v = 0.f;
    if (inv_logq) {
      float q = 0.f, lg = 0.f;
      for (int j = 0; j < d; ++j) {
        float var = th[dm.th_base_cov() + (int64_t)j * d + j];
        float u = xs[m * ldx + j] - th[dm.th_base_mean() + j];
        q += u * u / var;
        lg += logf(var);
      }
      v = 0.5f * q + 0.5f * lg + 0.91893853320467274f * (float)d;
    }
    ldr[m] = v;
  }
  __syncthreads();

  for (int step = 0; step < dm.L; ++step) {
    const int l = (a.dir == 0) ? step : dm.L - 1 - step;
    if (a.saved) {
      for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
        const int m = tw.m, j = tw.j;
        if (m0 + m < a.n) a.saved[((int64_t)l * a.n + m0 + m) * d + j] = xs[m * ldx + j];
      }
    }
    const float sc = th[dm.th_scale(l)], sh = th[dm.th_shift(l)];
    const bool affine_first = (a.dir != 2);
    // ---- ScalarAffine (forward / reference inverse apply it before the spline)
    if (affine_first) {
      const float e = expf(a.dir == 0 ? sc : -sc);
      for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
        const int m = tw.m, j = tw.j;
        float v = xs[m * ldx + j];
        xs[m * ldx + j] = (a.dir == 0) ? (v + sh) * e : v * e - sh;
      }
      for (int m = tid; m < MT; m += 256) ldr[m] += (a.dir == 0 ? sc : -sc) * (float)d;
      __syncthreads();
    }
    // ---- conditioner input (conditioning dims only)
    const int dc = dm.dc(l), kc = dm.kc(l);
    for (TileWalk tw(tid, kc); tw.m < MT; tw.next()) {
      const int m = tw.m, k = tw.j;
      a0[m * sm.la0 + k] = PT::cvt(k < dc ? xs[m * ldx + dm.cond(l, k)] : 0.f);
    }
    __syncthreads();
    const T* W = pk + dm.pk_layer(l);
    const float* bl = f.pbias + dm.pb_layer(l);
    const int h1p = dm.h1p(), h2p = dm.h2p();
    // ---- GEMM1: h1 = tanh(W1c c + b1)
    {
      const int KB = kc / PT::KSTEP, NBt = h1p / 32;
      for (int c0 = wave; c0 < NBt; c0 += 8) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PASS_PF, true>(a0, sm.la0, W + dm.pk_w1(l), KB, c0, 4, ncb, 0, KB, 0, acc, lane);
#pragma unroll
        for (int j = 0; j < 2; ++j)
          if (j < ncb) hidden_epilogue_tr<PT, MB>(hb1, sm.lh1, bl, c0 + 4 * j, acc[j], lane);
      }
    }
    __syncthreads();
    // ---- GEMM2: h2 = tanh(W2 h1 + b2)
    {
      const int KB = h1p / PT::KSTEP, NBt = h2p / 32;
      for (int c0 = wave; c0 < NBt; c0 += 8) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PASS_PF, true>(hb1, sm.lh1, W + dm.pk_w2(l), KB, c0, 4, ncb, 0, KB, 0, acc, lane);
#pragma unroll
        for (int j = 0; j < 2; ++j)
          if (j < ncb) hidden_epilogue_tr<PT, MB>(hb2, sm.lh2, bl + h1p, c0 + 4 * j, acc[j], lane);
      }
    }
    __syncthreads();
    // ---- GEMM3 (transposed, register-spline order) + spline in registers
    if (dm.reg_spline()) {
      const int npair = dm.n_pairs(l), KB = h2p / PT::KSTEP, dtl = dm.dt(l), K = dm.K;
      const T* W3r = W + dm.pk_w3r(l);
      const float* b3r = bl + dm.pb_b3r(l);
      float* ldw = ldp;  // [4 waves][MT] partial log-dets
      float ldacc[MB];
#pragma unroll
      for (int r = 0; r < MB; ++r) ldacc[r] = 0.f;
      const int hf = lane >> 5;
      for (int q = wave; q < npair; q += 4) {
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PASS_PF, true>(hb2, sm.lh2, W3r, KB, 2 * q, 1, 2, 0, KB, 0, acc, lane);
        const int k = 2 * q + hf;
        if (k < dtl) {
          const int j = dm.trans(l, k);
          const float4* bp = reinterpret_cast<const float4*>(b3r + q * 64 + hf * 32);
          float bias[32];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const float4 b4 = bp[u];
            bias[4 * u] = b4.x;
            bias[4 * u + 1] = b4.y;
            bias[4 * u + 2] = b4.z;
            bias[4 * u + 3] = b4.w;
          }
#pragma unroll
          for (int r = 0; r < MB; ++r) {
            float pp[32];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
              pp[v] = acc[0][r][v] + bias[v];
              pp[16 + v] = acc[1][r][v] + bias[16 + v];
            }
            const int m = r * 32 + (lane & 31);
            const float xv = xs[m * ldx + j];
            float yv, ldv;
            if (a.dir == 0)
              rs_fwd<KT>(pp, K, f, xv, yv, ldv);
            else
              rs_inv<KT>(pp, K, f, xv, yv, ldv);
            xs[m * ldx + j] = yv;
            ldacc[r] += ldv;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < MB; ++r) {
        const float o = __shfl_xor(ldacc[r], 32, 64);
        if (lane < 32) ldw[wave * MT + r * 32 + lane] = ldacc[r] + o;
      }
      __syncthreads();
      for (int m = tid; m < MT; m += 256) ldr[m] += (ldw[m] + ldw[MT + m]) + (ldw[2 * MT + m] + ldw[3 * MT + m]);
      __syncthreads();
    }
    // ---- GEMM3 in chunks of G transformed dims + spline epilogue (K > 10)
    const int nch = dm.reg_spline() ? 0 : dm.n_chunks(l), P = dm.P();
    for (int c = 0; c < nch; ++c) {
      const int nc = dm.chunk_nc(l, c), cnt = dm.chunk_cnt(l, c), coff = dm.chunk_off(l, c);
      const int KB = h2p / PT::KSTEP, NBt = nc / 32;
      for (int c0 = wave; c0 < NBt; c0 += 8) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PASS_PF>(hb2, sm.lh2, W + dm.pk_w3(l), KB, coff / 32 + c0, 4, ncb, 0, KB, 0, acc,
                                        lane);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < ncb) {
            const int col = (c0 + 4 * j) * 32 + (lane & 31);
            const float bb = bl[h1p + h2p + coff + col];
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v) obuf[(r * 32 + out_row(v, lane)) * sm.lo_ + col] = acc[j][r][v] + bb;
          }
        }
      }
      __syncthreads();
      for (int i = tid; i < MT * cnt; i += 256) {
        int m = i % MT, jj = i / MT;
        int j = dm.trans(l, c * dm.G() + jj);
        const float* pp = obuf + m * sm.lo_ + jj * P;
        float xv = xs[m * ldx + j], yv, ldv;
        if (a.dir == 0)
          spline_fwd(pp, f, xv, yv, ldv);
        else
          spline_inv(pp, f, xv, yv, ldv);
        xs[m * ldx + j] = yv;
        ldp[m * dm.G() + jj] = ldv;
      }
      __syncthreads();
      for (int m = tid; m < MT; m += 256) {
        float s = ldr[m];
        for (int jj = 0; jj < cnt; ++jj) s += ldp[m * dm.G() + jj];
        ldr[m] = s;
      }
      __syncthreads();
    }
    if (!affine_first) {  // exact inverse: affine^-1 after spline^-1
      const float e = expf(-sc);
      for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
        const int m = tw.m, j = tw.j;
        xs[m * ldx + j] = xs[m * ldx + j] * e - sh;
      }
      for (int m = tid; m < MT; m += 256) ldr[m] += -sc * (float)d;
      __syncthreads();
    }
  }
  // ---- outputs
  for (int m = tid; m < MT; m += 256) {
    int64_t row = m0 + m;
    if (row >= a.n) continue;
    float ld = ldr[m];
    if (a.dir == 0 && (a.flags & FL_BASE)) {
      // log N(y; base_mean, diag(base_cov)) (common.py:285-286 with a diagonal covariance)
      float q = 0.f, lg = 0.f;
      for (int j = 0; j < d; ++j) {
        float var = th[dm.th_base_cov() + (int64_t)j * d + j];
        float u = xs[m * ldx + j] - th[dm.th_base_mean() + j];
        q += u * u / var;
        lg += logf(var);
        if (a.gy_out) a.gy_out[row * d + j] = a.g_logp[row] * (-u / var);
      }
      ld += -0.5f * q - 0.5f * lg - 0.91893853320467274f * (float)d;
    }
    if (inv_logq) ld = -ld;
    if (a.logdet) a.logdet[row] = ld;
  }
  if (a.x_out) {
    for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
      const int m = tw.m, j = tw.j;
      int64_t row = m0 + m;
      if (row >= a.n) continue;
      float v = xs[m * ldx + j];
      if (a.dir != 0 && (a.flags & FL_STD))
        v = v * sqrtf(th[dm.th_data_cov() + (int64_t)j * d + j]) + th[dm.th_data_mean() + j];
      a.x_out[row * d + j] = v;
    }
  }
}

// ------------------------------------------------------ backward kernel --
struct BwdArgs {
  int layer;
  const float* x_in;     // [n, d] input of this layer (after standardisation)
  int64_t n;
  const float* g_logp;   // [n]
  const float* gy;       // [n, d] dL/d(layer output)
  float* gx;             // [n, d] dL/d(layer input)
  void* store;           // fragment-blocked activations / gradients (st_rows x ld elements)
  int64_t ld;
  float* aff_part;       // [n_blocks, 2]
};

template <class PT>
struct BwdSmem {
  int xs, gb, glp, hb1, hb2, ob, gpb, red, total;
  int ldx, lh1, lh2, lo_, lg;
  __host__ __device__ BwdSmem(const FlowDims& dm, int MT) {
    ldx = dm.d + 1;
    lh1 = dm.h1p() + PT::PAD;
    lh2 = dm.h2p() + PT::PAD;
    lo_ = dm.ncf() + 1;
    int kcm = dm.kc(1) > dm.kc(0) ? dm.kc(1) : dm.kc(0);
    // gradient staging: one chunk (chunked path) or 4 dim pairs of 64 rows
    const int gw = dm.reg_spline() ? 256 : dm.ncf();
    lg = (gw > kcm ? gw : kcm) + PT::PAD;
    const int tb = (int)sizeof(typename PT::T);
    int o = 0;
    xs = o;
    o += rup(MT * ldx * 4, 16);
    gb = o;
    o += rup(MT * ldx * 4, 16);
    glp = o;
    o += rup(MT * 4, 16);
    hb1 = o;
    o += rup(MT * lh1 * tb, 16);
    hb2 = o;
    o += rup(MT * lh2 * tb, 16);
    ob = o;  // spline parameters of a chunk (chunked path only)
    o += dm.reg_spline() ? 0 : rup(MT * lo_ * 4, 16);
    gpb = o;
    o += rup(MT * lg * tb, 16);
    red = o;
    o += 64;
    total = o;
  }
};

// Tile of the fragment-blocked store: rows row0 .. row0+nf (row0 and nf
// multiples of 32) x samples m0 .. m0+MT, from src[m * lds + f] (rows f >=
// nvalid are written as zeros). One thread per EPL-element lane chunk, so
// consecutive threads write consecutive chunks of a 64-lane block.
template <class PT>
__device__ __forceinline__ void store_fm(typename PT::T* st, int64_t ld, int row0, int nf, int nvalid,
                                         const typename PT::T* src, int lds, int64_t m0, int MT) {
  using T = typename PT::T;
  constexpr int EPL = PT::EPL, KS = PT::KSTEP;
  const int64_t ns = ld / KS;
  const int spt = MT / KS;                 // batch slices in the tile
  const int nchunk = (nf / 32) * spt * 64;
  for (int c = threadIdx.x; c < nchunk; c += 256) {
    const int lane = c & 63, sl = (c >> 6) % spt, rbl = (c >> 6) / spt;
    const int f = rbl * 32 + (lane & 31);
    const int mb = sl * KS + (lane >> 5) * EPL;
    T* dst = st + (((int64_t)(row0 / 32 + rbl) * ns + m0 / KS + sl) * 64 + lane) * EPL;
    if constexpr (EPL * sizeof(T) == 16) {  // one 16-byte store per lane chunk
      union {
        T t[EPL];
        uint4 u;
      } pk;
#pragma unroll
      for (int e = 0; e < EPL; ++e) pk.t[e] = (f < nvalid) ? src[(mb + e) * lds + f] : PT::cvt(0.f);
      *reinterpret_cast<uint4*>(dst) = pk.u;
    } else {
#pragma unroll
      for (int e = 0; e < EPL; ++e) dst[e] = (f < nvalid) ? src[(mb + e) * lds + f] : PT::cvt(0.f);
    }
  }
}

template <class PT, int MB, int HMAX, int KT>
__device__ __forceinline__ void flow_bwd_body(const FlowDev& f, const BwdArgs& a) {
  using T = typename PT::T;
  constexpr int MT = 32 * MB;
  // 32x32 tiles per wave of an [MT x hidden] block (hidden <= HMAX); the same
  // wave owns the same tiles in the forward recompute and in the adjoint, so
  // the FP32 tanh derivatives 1 - h^2 stay in its registers (the LDS copy of
  // h may be BF16, whose rounding would wreck 1 - h^2 near saturation).
  constexpr int NCW = (HMAX / 32) / 4;  // column blocks per wave (even)
  extern __shared__ __align__(16) unsigned char smem[];
  const FlowDims& dm = f.dm;
  BwdSmem<PT> sm(dm, MT);
  float* xs = reinterpret_cast<float*>(smem + sm.xs);
  float* gb = reinterpret_cast<float*>(smem + sm.gb);
  float* glp = reinterpret_cast<float*>(smem + sm.glp);
  T* hb1 = reinterpret_cast<T*>(smem + sm.hb1);
  T* hb2 = reinterpret_cast<T*>(smem + sm.hb2);
  float* obuf = reinterpret_cast<float*>(smem + sm.ob);
  T* gpb = reinterpret_cast<T*>(smem + sm.gpb);
  float* red = reinterpret_cast<float*>(smem + sm.red);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int d = dm.d, ldx = sm.ldx, l = a.layer;
  const int64_t m0 = (int64_t)blockIdx.x * MT;
  const float* th = f.theta;
  const T* W = reinterpret_cast<const T*>(f.packed) + dm.pk_layer(l);
  const float* bl = f.pbias + dm.pb_layer(l);
  T* st = reinterpret_cast<T*>(a.store);
  const int h1p = dm.h1p(), h2p = dm.h2p(), P = dm.P();
  const float sc = th[dm.th_scale(l)], sh = th[dm.th_shift(l)];
  const float es = expf(sc);

  // ---- load x (layer input), gy, g_logp; apply the affine part
  for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
    const int m = tw.m, j = tw.j;
    int64_t row = m0 + m;
    bool ok = row < a.n;
    xs[m * ldx + j] = ok ? (a.x_in[row * d + j] + sh) * es : 0.f;
    gb[m * ldx + j] = ok ? a.gy[row * d + j] : 0.f;
  }
  for (int m = tid; m < MT; m += 256) glp[m] = (m0 + m < a.n) ? a.g_logp[m0 + m] : 0.f;
  __syncthreads();
  // ---- recompute conditioner forward
  const int dc = dm.dc(l), kc = dm.kc(l);
  T* a0 = gpb;
  const int la0 = sm.lg;
  for (TileWalk tw(tid, kc); tw.m < MT; tw.next()) {
    const int m = tw.m, k = tw.j;
    a0[m * la0 + k] = PT::cvt(k < dc ? xs[m * ldx + dm.cond(l, k)] : 0.f);
  }
  __syncthreads();
  store_fm<PT>(st, a.ld, (int)dm.st_c(l), dm.ncp(l), kc, a0, la0, m0, MT);
  // wave w owns column blocks w, w+4, w+8, ... (all MB row blocks) of every
  // [MT x hidden] product below; index ci of the register arrays is the
  // wave-local column block.
  f32x16_t dh1[NCW][MB], dh2[NCW][MB];
  {
    const int KB = kc / PT::KSTEP, NBt = h1p / 32;
#pragma unroll
    for (int ci = 0; ci < NCW; ci += 2) {
      const int c0 = wave + 4 * ci;
      if (c0 < NBt) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PF_DEPTH>(a0, la0, W + dm.pk_w1(l), KB, c0, 4, ncb, 0, KB, 0, acc, lane);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < ncb) {
            const int col = (c0 + 4 * j) * 32 + (lane & 31);
            const float bb = bl[col];
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v) {
                float h = PT::act(acc[j][r][v] + bb);
                hb1[(r * 32 + out_row(v, lane)) * sm.lh1 + col] = PT::cvt(h);
                dh1[ci + j][r][v] = 1.f - h * h;
              }
          }
        }
      }
    }
  }
  __syncthreads();
  store_fm<PT>(st, a.ld, (int)dm.st_h1(l), h1p, h1p, hb1, sm.lh1, m0, MT);
  {
    const int KB = h1p / PT::KSTEP, NBt = h2p / 32;
#pragma unroll
    for (int ci = 0; ci < NCW; ci += 2) {
      const int c0 = wave + 4 * ci;
      if (c0 < NBt) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PF_DEPTH>(hb1, sm.lh1, W + dm.pk_w2(l), KB, c0, 4, ncb, 0, KB, 0, acc, lane);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < ncb) {
            const int col = (c0 + 4 * j) * 32 + (lane & 31);
            const float bb = bl[h1p + col];
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v) {
                float h = PT::act(acc[j][r][v] + bb);
                hb2[(r * 32 + out_row(v, lane)) * sm.lh2 + col] = PT::cvt(h);
                dh2[ci + j][r][v] = 1.f - h * h;
              }
          }
        }
      }
    }
  }
  __syncthreads();
  store_fm<PT>(st, a.ld, (int)dm.st_h2(l), h2p, h2p, hb2, sm.lh2, m0, MT);

  // ---- chunks: params -> spline VJP -> gp; g_h2 += gp W3c (register accumulators)
  f32x16_t gh2[NCW][MB];
#pragma unroll
  for (int ci = 0; ci < NCW; ++ci)
#pragma unroll
    for (int r = 0; r < MB; ++r) gh2[ci][r] = f32x16_t{};
  const int NB2 = h2p / 32;
  if (dm.reg_spline()) {
    // rounds of 4 dim pairs (one per wave): transposed GEMM3 -> parameters in
    // registers -> spline VJP in registers -> gp (gradient order) in LDS ->
    // gradient store + g_h2 += gp W3
    const int npair = dm.n_pairs(l), dtl = dm.dt(l), K = dm.K;
    const T* W3r = W + dm.pk_w3r(l);
    const T* W3rt = W + dm.pk_w3rt(l);
    const float* b3r = bl + dm.pb_b3r(l);
    const int KB3 = h2p / PT::KSTEP, KB3t = dm.n3r(l) / PT::KSTEP;
    const int hf = lane >> 5;
    for (int q0 = 0; q0 < npair; q0 += 4) {
      const int nq = (npair - q0) < 4 ? (npair - q0) : 4;
      if (wave < nq) {
        const int q = q0 + wave;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PF_DEPTH, true>(hb2, sm.lh2, W3r, KB3, 2 * q, 1, 2, 0, KB3, 0, acc, lane);
        const int k = 2 * q + hf;
        const bool valid = k < dtl;
        const int j = valid ? dm.trans(l, k) : 0;
        const float4* bp = reinterpret_cast<const float4*>(b3r + q * 64 + hf * 32);
        float bias[32];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float4 b4 = bp[u];
          bias[4 * u] = b4.x;
          bias[4 * u + 1] = b4.y;
          bias[4 * u + 2] = b4.z;
          bias[4 * u + 3] = b4.w;
        }
#pragma unroll
        for (int r = 0; r < MB; ++r) {
          float pp[32], go[32];
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            pp[v] = acc[0][r][v] + bias[v];
            pp[16 + v] = acc[1][r][v] + bias[16 + v];
          }
          const int m = r * 32 + (lane & 31);
          if (valid) {
            const float gxv = rs_vjp<KT>(pp, K, f, xs[m * ldx + j], gb[m * ldx + j], glp[m], go);
            gb[m * ldx + j] = gxv;
          } else {
#pragma unroll
            for (int i = 0; i < 32; ++i) go[i] = 0.f;
          }
          // slot 8t+e of this lane's dim -> gradient-order column 64 wave + 16t + 8hf + e
          T* dst = gpb + m * sm.lg + 64 * wave + 8 * hf;
#pragma unroll
          for (int t4 = 0; t4 < 4; ++t4) {
            if constexpr (sizeof(T) == 2) {  // 8 BF16 slots: one 16-byte LDS store
              union {
                T t[8];
                uint4 u;
              } pk;
#pragma unroll
              for (int e = 0; e < 8; ++e) pk.t[e] = PT::cvt(go[8 * t4 + e]);
              *reinterpret_cast<uint4*>(dst + 16 * t4) = pk.u;
            } else {
#pragma unroll
              for (int e = 0; e < 8; ++e) dst[16 * t4 + e] = PT::cvt(go[8 * t4 + e]);
            }
          }
        }
      }
      __syncthreads();
      store_fm<PT>(st, a.ld, (int)dm.st_gp(l) + 64 * q0, 64 * nq, 64 * nq, gpb, sm.lg, m0, MT);
      const int kb0 = 64 * q0 / PT::KSTEP, kb1 = 64 * (q0 + nq) / PT::KSTEP;
#pragma unroll
      for (int ci = 0; ci < NCW; ci += 2) {
        const int c0 = wave + 4 * ci;
        if (c0 < NB2) {
          const int ncb = (c0 + 4 < NB2) ? 2 : 1;
          f32x16_t acc[2][MB];
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < MB; ++r) acc[j][r] = gh2[ci + j][r];
          mfma_panel<PT, MB, 2, PF_DEPTH>(gpb, sm.lg, W3rt, KB3t, c0, 4, ncb, kb0, kb1, 0, acc, lane);
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < MB; ++r) gh2[ci + j][r] = acc[j][r];
        }
      }
      __syncthreads();
    }
  }
  const int nch = dm.reg_spline() ? 0 : dm.n_chunks(l);
  const int KB3 = h2p / PT::KSTEP;
  const int KB3t = dm.n3p(l) / PT::KSTEP;
  for (int c = 0; c < nch; ++c) {
    const int nc = dm.chunk_nc(l, c), cnt = dm.chunk_cnt(l, c), coff = dm.chunk_off(l, c);
    const int NBt = nc / 32;
    for (int c0 = wave; c0 < NBt; c0 += 8) {
      const int ncb = (c0 + 4 < NBt) ? 2 : 1;
      f32x16_t acc[2][MB] = {};
      mfma_panel<PT, MB, 2, PF_DEPTH>(hb2, sm.lh2, W + dm.pk_w3(l), KB3, coff / 32 + c0, 4, ncb, 0, KB3, 0, acc,
                                      lane);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (j < ncb) {
          const int col = (c0 + 4 * j) * 32 + (lane & 31);
          const float bb = bl[h1p + h2p + coff + col];
#pragma unroll
          for (int r = 0; r < MB; ++r)
#pragma unroll
            for (int v = 0; v < 16; ++v) obuf[(r * 32 + out_row(v, lane)) * sm.lo_ + col] = acc[j][r][v] + bb;
        }
      }
    }
    __syncthreads();
    for (int i = tid; i < MT * cnt; i += 256) {
      int m = i % MT, jj = i / MT;
      int j = dm.trans(l, c * dm.G() + jj);
      const float* pp = obuf + m * sm.lo_ + jj * P;
      float gyv = gb[m * ldx + j];
      float gxv = spline_vjp<T>(pp, f, xs[m * ldx + j], gyv, glp[m], gpb + m * sm.lg + jj * P,
                                [](float v) { return PT::cvt(v); });
      gb[m * ldx + j] = gxv;
    }
    for (int i = tid; i < MT * (nc - cnt * P); i += 256) {
      int w = nc - cnt * P;
      int m = i / w, cc = cnt * P + (i - m * w);
      gpb[m * sm.lg + cc] = PT::cvt(0.f);
    }
    __syncthreads();
    store_fm<PT>(st, a.ld, (int)dm.st_gp(l) + coff, nc, nc, gpb, sm.lg, m0, MT);
    const int kb0 = coff / PT::KSTEP, kb1 = (coff + nc) / PT::KSTEP;
#pragma unroll
    for (int ci = 0; ci < NCW; ci += 2) {
      const int c0 = wave + 4 * ci;
      if (c0 < NB2) {
        const int ncb = (c0 + 4 < NB2) ? 2 : 1;
        f32x16_t acc[2][MB];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < MB; ++r) acc[j][r] = gh2[ci + j][r];
        mfma_panel<PT, MB, 2, PF_DEPTH>(gpb, sm.lg, W + dm.pk_w3t(l), KB3t, c0, 4, ncb, kb0, kb1, 0, acc, lane);
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < MB; ++r) gh2[ci + j][r] = acc[j][r];
      }
    }
    __syncthreads();
  }
  // ---- g_pre2 = g_h2 * (1 - h2^2)   (in place of h2)
#pragma unroll
  for (int ci = 0; ci < NCW; ++ci) {
    const int nb = wave + 4 * ci;
    if (nb < NB2) {
      const int col = nb * 32 + (lane & 31);
#pragma unroll
      for (int r = 0; r < MB; ++r)
#pragma unroll
        for (int v = 0; v < 16; ++v)
          hb2[(r * 32 + out_row(v, lane)) * sm.lh2 + col] = PT::cvt(gh2[ci][r][v] * dh2[ci][r][v]);
    }
  }
  __syncthreads();
  store_fm<PT>(st, a.ld, (int)dm.st_gp2(l), h2p, h2p, hb2, sm.lh2, m0, MT);
  // ---- g_pre1 = (g_pre2 W2) * (1 - h1^2)   (in place of h1)
  {
    const int KB = h2p / PT::KSTEP, NBt = h1p / 32;
#pragma unroll
    for (int ci = 0; ci < NCW; ci += 2) {
      const int c0 = wave + 4 * ci;
      if (c0 < NBt) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PF_DEPTH>(hb2, sm.lh2, W + dm.pk_w2t(l), KB, c0, 4, ncb, 0, KB, 0, acc, lane);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < ncb) {
            const int col = (c0 + 4 * j) * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v)
                hb1[(r * 32 + out_row(v, lane)) * sm.lh1 + col] = PT::cvt(acc[j][r][v] * dh1[ci + j][r][v]);
          }
        }
      }
    }
  }
  __syncthreads();
  store_fm<PT>(st, a.ld, (int)dm.st_gp1(l), h1p, h1p, hb1, sm.lh1, m0, MT);
  // ---- g_c = g_pre1 W1c, scattered onto the conditioning dims
  {
    const int KB = h1p / PT::KSTEP, NBt = dm.ncp(l) / 32;
    for (int c0 = wave; c0 < NBt; c0 += 8) {
      const int ncb = (c0 + 4 < NBt) ? 2 : 1;
      f32x16_t acc[2][MB] = {};
      mfma_panel<PT, MB, 2, PF_DEPTH>(hb1, sm.lh1, W + dm.pk_w1t(l), KB, c0, 4, ncb, 0, KB, 0, acc, lane);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (j < ncb) {
          const int k = (c0 + 4 * j) * 32 + (lane & 31);
          if (k < dc) {
            const int jd = dm.cond(l, k);
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v) gb[(r * 32 + out_row(v, lane)) * ldx + jd] += acc[j][r][v];
          }
        }
      }
    }
  }
  __syncthreads();
  // ---- affine backward: xa = (x + shift) e^s, ld += d*s
  float ps = 0.f, pt = 0.f;
  for (TileWalk tw(tid, d); tw.m < MT; tw.next()) {
    const int m = tw.m, j = tw.j;
    float g = gb[m * ldx + j];
    ps += g * xs[m * ldx + j];
    pt += g * es;
    int64_t row = m0 + m;
    if (row < a.n) a.gx[row * d + j] = g * es;
  }
  for (int m = tid; m < MT; m += 256) ps += (float)d * glp[m];
  for (int o = 32; o > 0; o >>= 1) {
    ps += __shfl_xor(ps, o, 64);
    pt += __shfl_xor(pt, o, 64);
  }
  if (lane == 0) {
    red[2 * wave] = ps;
    red[2 * wave + 1] = pt;
  }
  __syncthreads();
  if (tid == 0) {
    float s0 = 0.f, s1 = 0.f;
    for (int w = 0; w < 4; ++w) {
      s0 += red[2 * w];
      s1 += red[2 * w + 1];
    }
    // partials at a fixed 32-sample granularity: slot MB*block holds the
    // block's sums, the other MB-1 slots of the block hold zeros
    a.aff_part[2 * MB * blockIdx.x] = s0;
    a.aff_part[2 * MB * blockIdx.x + 1] = s1;
#pragma unroll
    for (int r = 1; r < MB; ++r) {
      a.aff_part[2 * (MB * blockIdx.x + r)] = 0.f;
      a.aff_part[2 * (MB * blockIdx.x + r) + 1] = 0.f;
    }
  }
}

// 32-sample tiles. Hidden widths up to 256: register budget held to two
// waves per SIMD (no spills), and with ~68 KB of LDS per tile two workgroups
// share a CU. Wider hidden layers keep the default budget.
template <class PT, int KT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
flow_bwd_kernel_h256(FlowDev f, BwdArgs a) {
  flow_bwd_body<PT, 1, 256, KT>(f, a);
}

template <class PT, int KT>
__global__ void __launch_bounds__(256) flow_bwd_kernel_h512(FlowDev f, BwdArgs a) {
  flow_bwd_body<PT, 1, 512, KT>(f, a);
}

// compile-time bin count for the register spline: 8 and 10 (the common
// configurations), 0 = runtime K (any K <= 10, and the chunked path)
inline int spline_kt(int K) { return K == 8 ? 8 : (K == 10 ? 10 : 0); }

// --------------------------------------------------- weight gradients ----
// One workgroup of WG_SPLIT wavefronts per 32x32 tile of gW = G^T Act: the
// batch reduction is split over the wavefronts (each takes an interleaved
// share of the KSTEP-wide batch slices) and the partial tiles are summed in
// LDS in a fixed order. Both operands come from the fragment-blocked store:
// each fragment is one coalesced 64-lane read.
// Problems: 0 = W1c, 1 = W2, 2 = W3 (gradient-store order). Tiles of the
// first column block (kb == 0) also sum their G rows into the bias
// gradients. The last tile sums the ScalarAffine partials. Workgroup b runs
// on XCD b % 8; tiles are dealt so that XCD x gets a contiguous range of
// them, i.e. the tiles sharing a G row panel share one L2.
constexpr int WG_SPLIT = 8;
static_assert(16 % WG_SPLIT == 0, "each wave finishes a whole number of accumulator registers");
constexpr int WG_XCDS = 8;

__host__ __device__ inline int wgrad_ntiles(const FlowDims& dm, int l) {
  return (dm.h1p() / 32) * (dm.ncp(l) / 32) + (dm.h2p() / 32) * (dm.h1p() / 32) + (dm.n3(l) / 32) * (dm.h2p() / 32) +
         1;
}

template <class PT>
__global__ void __launch_bounds__(64 * WG_SPLIT) flow_wgrad_kernel(FlowDev f, int l, const void* store, int64_t ld,
                                                                   const float* aff_part, int n_part, float* grad) {
  using T = typename PT::T;
  __shared__ float red[WG_SPLIT][16][64];
  __shared__ float rsum[WG_SPLIT][64];
  const FlowDims& dm = f.dm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int total = wgrad_ntiles(dm, l);
  const int per = (total + WG_XCDS - 1) / WG_XCDS;
  int tt = (int)(blockIdx.x % WG_XCDS) * per + (int)(blockIdx.x / WG_XCDS);
  if (tt >= total) return;
  const int h1p = dm.h1p(), h2p = dm.h2p();
  const int n0t = (h1p / 32) * (dm.ncp(l) / 32), n1t = (h2p / 32) * (h1p / 32), n2t = (dm.n3(l) / 32) * (h2p / 32);
  if (tt == n0t + n1t + n2t) {  // ScalarAffine gradients: fixed-order sum of the partials
    if (threadIdx.x == 0) {
      float s0 = 0.f, s1 = 0.f;
      for (int i = 0; i < n_part; ++i) {
        s0 += aff_part[2 * i];
        s1 += aff_part[2 * i + 1];
      }
      grad[dm.th_scale(l)] = s0;
      grad[dm.th_shift(l)] = s1;
    }
    return;
  }
  int prob;
  if (tt < n0t) {
    prob = 0;
  } else if ((tt -= n0t) < n1t) {
    prob = 1;
  } else {
    tt -= n1t;
    prob = 2;
  }
  int gRow, aRow, ncolb;
  if (prob == 0) {
    gRow = (int)dm.st_gp1(l);
    aRow = (int)dm.st_c(l);
    ncolb = dm.ncp(l) / 32;
  } else if (prob == 1) {
    gRow = (int)dm.st_gp2(l);
    aRow = (int)dm.st_h1(l);
    ncolb = h1p / 32;
  } else {
    gRow = (int)dm.st_gp(l);
    aRow = (int)dm.st_h2(l);
    ncolb = h2p / 32;
  }
  const int nb = tt / ncolb, kb = tt - nb * ncolb;
  const bool want_bias = (kb == 0);
  const T* st = reinterpret_cast<const T*>(store);
  f32x16_t acc = {};
  float bsum = 0.f;
  // this wave takes batch slices wave, wave + WG_SPLIT, ... ; fragments
  // (L2-resident store) are prefetched P slices ahead
  constexpr int P = PF_DEPTH;
  using GF = decltype(PT::load_g(st, 0, 0, 0, 0));
  GF ga[P], ab[P];
  const int nsteps = (int)(ld / PT::KSTEP);
  const int gRb = gRow / 32 + nb, aRb = aRow / 32 + kb;
  const int my = (nsteps - wave + WG_SPLIT - 1) / WG_SPLIT;  // slices of this wave
#pragma unroll
  for (int p = 0; p < P; ++p)
    if (p < my) {
      const int sl = wave + p * WG_SPLIT;
      ga[p] = PT::load_g(st, nsteps, gRb, sl, lane);
      ab[p] = PT::load_g(st, nsteps, aRb, sl, lane);
    }
  for (int s0 = 0; s0 < my; s0 += P) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int i = s0 + p;
      if (i < my) {
        acc = PT::mma(ga[p], ab[p], acc);
        if (want_bias) bsum += PT::hsum(ga[p]);
        if (i + P < my) {
          const int sl = wave + (i + P) * WG_SPLIT;
          ga[p] = PT::load_g(st, nsteps, gRb, sl, lane);
          ab[p] = PT::load_g(st, nsteps, aRb, sl, lane);
        }
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) red[wave][v][lane] = acc[v];
  rsum[wave][lane] = bsum;
  __syncthreads();
  // each wave finishes 16 / WG_SPLIT accumulator registers (fixed-order sum
  // over the waves' partials); the last wave also finishes the bias row
  constexpr int VPW = 16 / WG_SPLIT;
  const int k = kb * 32 + (lane & 31);
#pragma unroll
  for (int u = 0; u < VPW; ++u) {
    const int v = wave * VPW + u;
    float t = 0.f;
#pragma unroll
    for (int w = 0; w < WG_SPLIT; ++w) t += red[w][v][lane];
    const int n = nb * 32 + out_row(v, lane);
    int64_t idx = -1;
    if (prob == 0) {
      if (n < dm.h1 && k < dm.dc(l)) idx = dm.th_w1(l) + (int64_t)n * dm.d + dm.cond(l, k);
    } else if (prob == 1) {
      if (n < dm.h2 && k < dm.h1) idx = dm.th_w2(l) + (int64_t)n * dm.h1 + k;
    } else {
      int row = dm.gp_row(l, n);
      if (row >= 0 && k < dm.h2) idx = dm.th_w3(l) + (int64_t)row * dm.h2 + k;
    }
    if (idx >= 0) grad[idx] = t;
  }
  if (want_bias && wave == WG_SPLIT - 1) {
    // lane l and l+32 hold the two halves of G row nb*32 + l%32
    float b = 0.f;
#pragma unroll
    for (int w = 0; w < WG_SPLIT; ++w) b += rsum[w][lane];
    b += __shfl_xor(b, 32, 64);
    const int n = nb * 32 + lane;
    if (lane < 32) {
      int64_t idx = -1;
      if (prob == 0) {
        if (n < dm.h1) idx = dm.th_b1(l) + n;
      } else if (prob == 1) {
        if (n < dm.h2) idx = dm.th_b2(l) + n;
      } else {
        int row = dm.gp_row(l, n);
        if (row >= 0) idx = dm.th_b3(l) + row;
      }
      if (idx >= 0) grad[idx] = b;
    }
  }
}

// ------------------------------------------------------------- packing ---
// Grid: x strides over the elements of one image, y = layer * 9 + image, with
// images 0..7 = w1, w2, w3, w1t, w2t, w3t, w3r, w3rt and 8 = the layer's
// packed biases, so every element decodes its own (block, lane, k) only.
template <class PT>
__global__ void __launch_bounds__(256) flow_pack_kernel(FlowDev f, typename PT::T* packed, float* pbias) {
  const FlowDims& dm = f.dm;
  const float* th = f.theta;
  const int l = blockIdx.y / 9, which = blockIdx.y - 9 * l;
  const int h1p = dm.h1p(), h2p = dm.h2p();
  if (which == 8) {
    const int64_t n = dm.pb_size(l);
    float* out = pbias + dm.pb_layer(l);
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
      float v = 0.f;
      if (e < h1p) {
        if (e < dm.h1) v = th[dm.th_b1(l) + e];
      } else if (e < h1p + h2p) {
        int q = (int)(e - h1p);
        if (q < dm.h2) v = th[dm.th_b2(l) + q];
      } else if (e < dm.pb_b3r(l)) {
        int row = dm.w3_row(l, (int)(e - h1p - h2p));
        if (row >= 0) v = th[dm.th_b3(l) + row];
      } else {
        int row = dm.b3r_row(l, (int)(e - dm.pb_b3r(l)));
        if (row >= 0) v = th[dm.th_b3(l) + row];
      }
      out[e] = v;
    }
    return;
  }
  int64_t off;
  int N, Kdim;
  switch (which) {
    case 0: off = dm.pk_w1(l); N = h1p; Kdim = dm.kc(l); break;
    case 1: off = dm.pk_w2(l); N = h2p; Kdim = h1p; break;
    case 2: off = dm.pk_w3(l); N = dm.n3p(l); Kdim = h2p; break;
    case 3: off = dm.pk_w1t(l); N = dm.ncp(l); Kdim = h1p; break;
    case 4: off = dm.pk_w2t(l); N = h1p; Kdim = h2p; break;
    case 5: off = dm.pk_w3t(l); N = h2p; Kdim = dm.n3p(l); break;
    case 6: off = dm.pk_w3r(l); N = dm.n3r(l); Kdim = h2p; break;
    default: off = dm.pk_w3rt(l); N = h2p; Kdim = dm.n3r(l); break;
  }
  const int64_t size = (int64_t)N * Kdim;
  typename PT::T* out = packed + dm.pk_layer(l) + off;
  const int KB = Kdim / PT::KSTEP;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < size; idx += (int64_t)gridDim.x * 256) {
    const int ee = (int)(idx % PT::EPL);
    const int lane = (int)((idx / PT::EPL) % 64);
    const int64_t blk = idx / (64 * PT::EPL);
    const int kb = (int)(blk % KB), nb = (int)(blk / KB);
    const int n = nb * 32 + (lane & 31);
    const int k = kb * PT::KSTEP + (lane >> 5) * PT::EPL + ee;
    float v = 0.f;
    switch (which) {
      case 0:
        if (n < dm.h1 && k < dm.dc(l)) v = th[dm.th_w1(l) + (int64_t)n * dm.d + dm.cond(l, k)];
        break;
      case 1:
        if (n < dm.h2 && k < dm.h1) v = th[dm.th_w2(l) + (int64_t)n * dm.h1 + k];
        break;
      case 2: {
        int row = dm.w3_row(l, n);
        if (row >= 0 && k < dm.h2) v = th[dm.th_w3(l) + (int64_t)row * dm.h2 + k];
      } break;
      case 3:
        if (n < dm.dc(l) && k < dm.h1) v = th[dm.th_w1(l) + (int64_t)k * dm.d + dm.cond(l, n)];
        break;
      case 4:
        if (n < dm.h1 && k < dm.h2) v = th[dm.th_w2(l) + (int64_t)k * dm.h1 + n];
        break;
      case 5: {
        int row = dm.w3_row(l, k);
        if (row >= 0 && n < dm.h2) v = th[dm.th_w3(l) + (int64_t)row * dm.h2 + n];
      } break;
      case 6: {
        int row = dm.w3r_row(l, n);
        if (row >= 0 && k < dm.h2) v = th[dm.th_w3(l) + (int64_t)row * dm.h2 + k];
      } break;
      default: {
        int row = dm.w3k_row(l, k);
        if (row >= 0 && n < dm.h2) v = th[dm.th_w3(l) + (int64_t)row * dm.h2 + n];
      } break;
    }
    out[idx] = PT::cvt(v);
  }
}

// --------------------------------------------------------------- AdamW ---
__global__ void __launch_bounds__(256) sumsq_kernel(const float* __restrict__ g, int64_t n, double* part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v = g[i];
    s += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double r[4];
  if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = r[0] + r[1] + r[2] + r[3];
}

struct AdamHyper {
  float lr, b1, b2, eps, wd, max_norm;
  float bc1, bc2;  // 1 - b1^t, 1 - b2^t
};

// Global norm: fixed-order tree sum of the per-block partials, one block.
// sqrt of the sum goes to part[n_part] (and, as float, to norm_out).
__global__ void __launch_bounds__(256) norm_kernel(double* part, int n_part, float* norm_out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n_part; i += 256) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double r[4];
  if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double nrm = sqrt((r[0] + r[1]) + (r[2] + r[3]));
    part[n_part] = nrm;
    if (norm_out) *norm_out = (float)nrm;
  }
}

// optax.chain(clip_by_global_norm(max_norm), adamw(lr, b1, b2, eps, wd)):
// u = g (or g / |g| * max_norm when |g| >= max_norm); m, v moments;
// theta += -lr * (m_hat / (sqrt(v_hat) + eps) + wd * theta)
__device__ __forceinline__ void adamw_one(float& th, float g, float& mi, float& vi, const AdamHyper& h, bool clip,
                                          float gn) {
  float u = g;
  if (clip) u = (u / gn) * h.max_norm;
  mi = (1.f - h.b1) * u + h.b1 * mi;
  vi = (1.f - h.b2) * u * u + h.b2 * vi;
  const float mh = mi / h.bc1, vh = vi / h.bc2;
  const float up = mh / (sqrtf(vh) + h.eps) + h.wd * th;
  th = th + (-h.lr) * up;
}

__global__ void __launch_bounds__(256) adamw_kernel(float* theta, const float* __restrict__ g, float* m, float* v,
                                                    int64_t n, AdamHyper h, const double* nrm) {
  const float gn = (float)nrm[0];
  const bool clip = !(gn < h.max_norm);
  const int64_t n4 = n / 4;
  float4* th4 = reinterpret_cast<float4*>(theta);
  const float4* g4 = reinterpret_cast<const float4*>(g);
  float4* m4 = reinterpret_cast<float4*>(m);
  float4* v4 = reinterpret_cast<float4*>(v);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)gridDim.x * 256) {
    float4 t = th4[i], gg = g4[i], mm = m4[i], vv = v4[i];
    adamw_one(t.x, gg.x, mm.x, vv.x, h, clip, gn);
    adamw_one(t.y, gg.y, mm.y, vv.y, h, clip, gn);
    adamw_one(t.z, gg.z, mm.z, vv.z, h, clip, gn);
    adamw_one(t.w, gg.w, mm.w, vv.w, h, clip, gn);
    th4[i] = t;
    m4[i] = mm;
    v4[i] = vv;
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    float t = theta[i], mm = m[i], vv = v[i];
    adamw_one(t, g[i], mm, vv, h, clip, gn);
    theta[i] = t;
    m[i] = mm;
    v[i] = vv;
  }
}

// --------------------------------------------------------------- host ----
static int check_flow(const fmc_flow* fl, FlowDev& fd) {
  FMC_ARG(fl != nullptr, "flow is NULL");
  FMC_ARG(fl->n_dim >= 2 && fl->n_dim <= 512, "flow n_dim must be in [2, 512] (got %d)", fl->n_dim);
  FMC_ARG(fl->n_layers >= 1 && fl->n_layers <= 256, "flow n_layers must be in [1, 256]");
  FMC_ARG(fl->n_bins >= 1 && fl->n_bins <= 42, "flow n_bins must be in [1, 42]");
  FMC_ARG(fl->hidden1 >= 1 && fl->hidden1 <= 512 && fl->hidden2 >= 1 && fl->hidden2 <= 512,
          "hidden sizes must be in [1, 512]");
  FMC_ARG(fl->precision == FMC_PREC_F32 || fl->precision == FMC_PREC_BF16, "unknown precision %d", fl->precision);
  FMC_ARG(fl->theta != nullptr, "flow theta is NULL");
  FMC_ARG(fl->range_max > fl->range_min, "spline range must be increasing");
  FMC_ARG(fl->n_bins * fl->min_bin_size < fl->range_max - fl->range_min, "min_bin_size too large for the range");
  FMC_ARG(fl->min_knot_slope > 0.f && fl->min_knot_slope < 1.f, "min_knot_slope must be in (0, 1)");
  fd.dm = FlowDims{fl->n_dim, fl->n_layers, fl->n_bins, fl->hidden1, fl->hidden2,
                   fl->precision == FMC_PREC_BF16 ? 16 : 2};
  fd.theta = fl->theta;
  fd.packed = fl->packed;
  fd.pbias = fl->packed_bias;
  fd.lo = fl->range_min;
  fd.hi = fl->range_max;
  fd.mb = fl->min_bin_size;
  fd.ms = fl->min_knot_slope;
  fd.off = (float)log(exp(1.0 - (double)fl->min_knot_slope) - 1.0);
  return FMC_OK;
}

static int pass_mb(const FlowDims& dm, bool bf16, size_t& shm) {
  // 32-sample tiles (several workgroups per CU), if they fit the LDS budget
  int tot = bf16 ? PassSmem<PBF16>(dm, 32).total : PassSmem<PF32>(dm, 32).total;
  if (tot > 160 * 1024) return 0;
  shm = (size_t)tot;
  return 1;
}

static int bwd_mb(const FlowDims& dm, bool bf16, size_t& shm) {
  // 32-sample tiles, if they fit the LDS budget
  int tot = bf16 ? BwdSmem<PBF16>(dm, 32).total : BwdSmem<PF32>(dm, 32).total;
  if (tot > 160 * 1024) return 0;
  shm = (size_t)tot;
  return 1;
}

template <class K>
static void set_lds(K kernel, size_t shm) {
  hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
}

}  // namespace fmc

using namespace fmc;

extern "C" int fmc_flow_sizes(const fmc_flow* fl, int64_t* theta_size, int64_t* packed_size,
                              int64_t* packed_bias_size, int64_t* store_rows, int32_t* tile_fwd, int32_t* tile_bwd) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  const FlowDims& dm = fd.dm;
  if (theta_size) *theta_size = dm.th_total();
  if (packed_size) *packed_size = dm.pk_total();
  if (packed_bias_size) *packed_bias_size = dm.pb_total();
  if (store_rows) *store_rows = dm.st_rows(0) > dm.st_rows(1) ? dm.st_rows(0) : dm.st_rows(1);
  size_t s1 = 0, s2 = 0;
  int mbf = pass_mb(dm, fl->precision == FMC_PREC_BF16, s1);
  int mbb = bwd_mb(dm, fl->precision == FMC_PREC_BF16, s2);
  if (tile_fwd) *tile_fwd = 32 * mbf;
  // affine partial gradients are kept per 32 samples whatever tile the backward uses
  if (tile_bwd) *tile_bwd = mbb > 0 ? 32 : 0;
  return FMC_OK;
}

extern "C" int fmc_flow_pack(void* stream, const fmc_flow* fl) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(fl->packed != nullptr && fl->packed_bias != nullptr, "packed buffers are NULL");
  const FlowDims& dm = fd.dm;
  int64_t mx = 0;  // largest image of a layer
  for (int l = 0; l < (dm.L < 2 ? dm.L : 2); ++l) {
    const int64_t c[5] = {(int64_t)dm.h2p() * dm.h1p(), (int64_t)dm.n3p(l) * dm.h2p(), (int64_t)dm.n3r(l) * dm.h2p(),
                          (int64_t)dm.h1p() * dm.kc(l), dm.pb_size(l)};
    for (int64_t v : c) mx = v > mx ? v : mx;
  }
  unsigned gx = (unsigned)((mx + 255) / 256 < 128 ? (mx + 255) / 256 : 128);
  dim3 grid(gx < 1 ? 1 : gx, (unsigned)(9 * dm.L));
  hipStream_t st = (hipStream_t)stream;
  if (fl->precision == FMC_PREC_BF16)
    hipLaunchKernelGGL(flow_pack_kernel<PBF16>, grid, dim3(256), 0, st, fd, (__bf16*)fl->packed, fl->packed_bias);
  else
    hipLaunchKernelGGL(flow_pack_kernel<PF32>, grid, dim3(256), 0, st, fd, (float*)fl->packed, fl->packed_bias);
  return check_launch("flow pack");
}

extern "C" int fmc_flow_pass(void* stream, const fmc_flow* fl, int32_t direction, int32_t flags, const float* x_in,
                             int64_t n, uint32_t key0, uint32_t key1, int64_t sample0, float* x_out, float* logdet,
                             float* saved, const float* g_logp, float* gy_out) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(fl->packed != nullptr && fl->packed_bias != nullptr, "packed buffers are NULL (call fmc_flow_pack)");
  FMC_ARG(direction >= 0 && direction <= 2, "direction must be 0 (forward), 1 (inverse) or 2 (exact inverse)");
  FMC_ARG(n >= 0, "n < 0");
  FMC_ARG(x_in != nullptr || (direction != 0 && (flags & FL_SAMPLE)), "x_in is NULL");
  FMC_ARG(gy_out == nullptr || (g_logp != nullptr && direction == 0 && (flags & FL_BASE)),
          "gy_out needs g_logp, the forward direction and the base log-prob flag");
  FMC_ARG(saved == nullptr || direction == 0, "saved layer inputs are only produced by the forward direction");
  FMC_ARG(!(flags & FL_BASE) || direction != 1,
          "the base log-prob flag needs the forward or the exact inverse direction (the reference's inverse "
          "order is not the inverse of the forward map)");
  FMC_ARG(sample0 >= 0 && sample0 + n <= 0xFFFFFFFFLL, "sample index range exceeds 32 bits");
  if (n == 0) return FMC_OK;
  size_t shm = 0;
  const bool bf = fl->precision == FMC_PREC_BF16;
  int mb = pass_mb(fd.dm, bf, shm);
  FMC_REQUIRE(mb > 0, FMC_ERR_UNSUPPORTED, "flow dimensions need more than 160 KB of LDS per tile");
  PassArgs a{direction, flags, x_in, n, key0, key1, sample0, x_out, logdet, saved, g_logp, gy_out};
  const int MT = 32 * mb;
  dim3 g((unsigned)((n + MT - 1) / MT)), b(256);
  hipStream_t st = (hipStream_t)stream;
#define FMC_PASS(PT_, KT_)                                                 \
  do {                                                                     \
    set_lds(flow_pass_kernel<PT_, 1, KT_>, shm);                           \
    hipLaunchKernelGGL((flow_pass_kernel<PT_, 1, KT_>), g, b, shm, st, fd, a); \
  } while (0)
  const int kt = spline_kt(fd.dm.K);
  if (bf) {
    if (kt == 8) FMC_PASS(PBF16, 8); else if (kt == 10) FMC_PASS(PBF16, 10); else FMC_PASS(PBF16, 0);
  } else {
    if (kt == 8) FMC_PASS(PF32, 8); else if (kt == 10) FMC_PASS(PF32, 10); else FMC_PASS(PF32, 0);
  }
#undef FMC_PASS
  return check_launch("flow pass");
}

extern "C" int fmc_flow_backward_layer(void* stream, const fmc_flow* fl, int32_t layer, const float* x_layer,
                                       int64_t n, const float* g_logp, const float* gy, float* gx, void* store,
                                       int64_t store_ld, float* affine_partials, int32_t n_partials) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(fl->packed != nullptr && fl->packed_bias != nullptr, "packed buffers are NULL (call fmc_flow_pack)");
  FMC_ARG(layer >= 0 && layer < fl->n_layers, "layer out of range");
  FMC_ARG(x_layer && g_logp && gy && gx && store && affine_partials, "NULL pointer");
  FMC_ARG(n >= 1, "n must be >= 1");
  FMC_ARG(store_ld >= n && store_ld % 64 == 0, "store_ld must be a multiple of 64 and >= n");
  size_t shm = 0;
  const bool bf = fl->precision == FMC_PREC_BF16;
  int mb = bwd_mb(fd.dm, bf, shm);
  FMC_REQUIRE(mb > 0, FMC_ERR_UNSUPPORTED, "flow dimensions need more than 160 KB of LDS per tile");
  const int hmax = fd.dm.h1p() > fd.dm.h2p() ? fd.dm.h1p() : fd.dm.h2p();
  FMC_REQUIRE(hmax <= 512, FMC_ERR_UNSUPPORTED, "hidden sizes > 512 not supported by the backward kernel");
  // every store column [0, store_ld) is rewritten (rows >= n carry zero gradients)
  const int64_t nblk = store_ld / 32;
  FMC_ARG(n_partials >= nblk, "affine_partials holds %d slots, need %lld", n_partials, (long long)nblk);
  BwdArgs a{layer, x_layer, n, g_logp, gy, gx, store, store_ld, affine_partials};
  dim3 g((unsigned)nblk), b(256);
  hipStream_t st = (hipStream_t)stream;
#define FMC_BWD(KERN, PT_, KT_)                                     \
  do {                                                              \
    set_lds(KERN<PT_, KT_>, shm);                                   \
    hipLaunchKernelGGL((KERN<PT_, KT_>), g, b, shm, st, fd, a);     \
  } while (0)
#define FMC_BWD_K(KERN, PT_)                                        \
  do {                                                              \
    if (kt == 8) FMC_BWD(KERN, PT_, 8);                             \
    else if (kt == 10) FMC_BWD(KERN, PT_, 10);                      \
    else FMC_BWD(KERN, PT_, 0);                                     \
  } while (0)
  const int kt = spline_kt(fd.dm.K);
  if (hmax <= 256) {
    if (bf) FMC_BWD_K(flow_bwd_kernel_h256, PBF16); else FMC_BWD_K(flow_bwd_kernel_h256, PF32);
  } else {
    if (bf) FMC_BWD_K(flow_bwd_kernel_h512, PBF16); else FMC_BWD_K(flow_bwd_kernel_h512, PF32);
  }
#undef FMC_BWD_K
#undef FMC_BWD
  return check_launch("flow backward");
}

extern "C" int fmc_flow_weight_grad(void* stream, const fmc_flow* fl, int32_t layer, const void* store,
                                    int64_t store_ld, const float* affine_partials, int32_t n_partials, float* grad) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(layer >= 0 && layer < fl->n_layers, "layer out of range");
  FMC_ARG(store && affine_partials && grad, "NULL pointer");
  FMC_ARG(store_ld >= 64 && store_ld % 64 == 0, "store_ld must be a positive multiple of 64");
  const int total = wgrad_ntiles(fd.dm, layer);
  hipStream_t st = (hipStream_t)stream;
  const bool bf = fl->precision == FMC_PREC_BF16;
  FMC_ARG(n_partials >= store_ld / 32, "affine_partials holds %d slots, need %lld", n_partials,
          (long long)(store_ld / 32));
  const int per = (total + WG_XCDS - 1) / WG_XCDS;
  dim3 g((unsigned)(per * WG_XCDS)), b(64 * WG_SPLIT);
  if (bf)
    hipLaunchKernelGGL(flow_wgrad_kernel<PBF16>, g, b, 0, st, fd, layer, store, store_ld, affine_partials,
                       (int)(store_ld / 32), grad);
  else
    hipLaunchKernelGGL(flow_wgrad_kernel<PF32>, g, b, 0, st, fd, layer, store, store_ld, affine_partials,
                       (int)(store_ld / 32), grad);
  return check_launch("flow weight grad");
}

extern "C" int fmc_adamw_clip(void* stream, float* theta, const float* grad, float* m, float* v, int64_t n, float lr,
                              float b1, float b2, float eps, float weight_decay, float max_norm, int32_t step,
                              double* workspace, int32_t workspace_len, float* grad_norm) {
  FMC_ARG(theta && grad && m && v && workspace, "NULL pointer");
  FMC_ARG(n >= 0 && step >= 1, "n must be >= 0 and step >= 1");
  FMC_ARG(workspace_len >= 2, "workspace_len must be >= 2");
  FMC_ARG(((uintptr_t)theta | (uintptr_t)grad | (uintptr_t)m | (uintptr_t)v) % 16 == 0,
          "theta, grad, m and v must be 16-byte aligned");
  if (n == 0) return FMC_OK;
  int nb = (int)((n + 255) / 256);
  if (nb > workspace_len - 1) nb = workspace_len - 1;
  if (nb > 1024) nb = 1024;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(sumsq_kernel, dim3(nb), dim3(256), 0, st, grad, n, workspace);
  int rc = check_launch("adamw sumsq");
  if (rc) return rc;
  hipLaunchKernelGGL(norm_kernel, dim3(1), dim3(256), 0, st, workspace, nb, grad_norm);
  if ((rc = check_launch("adamw norm"))) return rc;
  AdamHyper h{lr, b1, b2, eps, weight_decay, max_norm, (float)(1.0 - pow((double)b1, step)),
              (float)(1.0 - pow((double)b2, step))};
  int nb2 = (int)((n / 4 + 255) / 256);
  if (nb2 > 2048) nb2 = 2048;
  if (nb2 < 1) nb2 = 1;
  hipLaunchKernelGGL(adamw_kernel, dim3(nb2), dim3(256), 0, st, theta, grad, m, v, n, h, workspace + nb);
  return check_launch("adamw update");
}
