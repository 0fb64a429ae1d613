// Shared device code and host helpers of the MaskedCouplingRQSpline flow
// kernels (fmc_flow.hip, fmc_flow_pass.hip, fmc_flow_bwd.hip): MFMA operand
// traits and panels, the rational-quadratic spline (LDS and register forms,
// forward / inverse / vector-Jacobian product), the pass and adjoint argument
// and LDS layouts, and the descriptor check.
//
// Reference: src/flowMC/resource/model/nf_model/rqSpline.py (spline :20-239,
// RQSpline :242-361, MaskedCouplingRQSpline :364-504), common.py (MLP
// :68-124, MaskedCouplingLayer :127-168, ScalarAffine :211-240, Gaussian
// :243-293).
#pragma once

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
  using BF = float;  // one B fragment (per lane)
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
  static __device__ __forceinline__ void act2(float a, float b, float& ha, float& hb) {
    ha = tanhf(a);
    hb = tanhf(b);
  }
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
  using BF = bf16x8_t;
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
  // two activations with one reciprocal: 1/da = r db, 1/db = r da with
  // r = 1/(da db) (Newton-refined); arguments clamped to |x| <= 15, where
  // tanh is 1 to FP32 precision, so da db <= (e^30 + 1)^2 stays finite
  static __device__ __forceinline__ void act2(float a, float b, float& ha, float& hb) {
    a = fminf(fmaxf(a, -15.f), 15.f);
    b = fminf(fmaxf(b, -15.f), 15.f);
    const float da = __expf(2.f * a) + 1.f, db = __expf(2.f * b) + 1.f;
    const float p = da * db;
    float r = __builtin_amdgcn_rcpf(p);
    r = fmaf(fmaf(-p, r, 1.f), r, r);
    ha = fmaf(-2.f, r * db, 1.f);
    hb = fmaf(-2.f, r * da, 1.f);
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
    // bias == nullptr: the accumulators started from the bias (bias_tr)
    const float4 b4 = bias ? *reinterpret_cast<const float4*>(bias + f0 + 8 * g) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int r = 0; r < MB; ++r)
    {
      float h0, h1, h2, h3;
      if (bias) {
        PT::act2(acc[r][4 * g] + b4.x, acc[r][4 * g + 1] + b4.y, h0, h1);
        PT::act2(acc[r][4 * g + 2] + b4.z, acc[r][4 * g + 3] + b4.w, h2, h3);
      } else {
        PT::act2(acc[r][4 * g], acc[r][4 * g + 1], h0, h1);
        PT::act2(acc[r][4 * g + 2], acc[r][4 * g + 3], h2, h3);
      }
      PT::store4(hb + (r * 32 + (lane & 31)) * lh + f0 + 8 * g, h0, h1, h2, h3);
    }
  }
}

// The bias of the transposed hidden tile (features fb*32.., the lane's rows
// as hidden_epilogue_tr reads them) as the accumulators' starting value: the
// MFMAs then add the products to it, which saves the epilogue's additions.
template <int MB>
__device__ __forceinline__ void bias_tr(f32x16_t (&acc)[MB], const float* bias, int fb, int lane) {
  const int f0 = fb * 32 + 4 * (lane >> 5);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 b4 = *reinterpret_cast<const float4*>(bias + f0 + 8 * g);
#pragma unroll
    for (int r = 0; r < MB; ++r) {
      acc[r][4 * g] = b4.x;
      acc[r][4 * g + 1] = b4.y;
      acc[r][4 * g + 2] = b4.z;
      acc[r][4 * g + 3] = b4.w;
    }
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

// Natural log and square root on the hardware instructions (v_log_f32 is
// log2 to about 1 ulp; v_sqrt_f32 about 1 ulp): every argument in the spline
// is a positive normal number (bin widths and heights >= the minimum bin
// size, slopes >= the minimum knot slope, 1 + e^-|v| in [1, 2], the
// discriminant clamped to FLT_MIN), where the library forms only add
// denormal scaling and extra-precision steps (12 and 16 instructions).
// FL = true for the BF16 conditioner; the FP32 conditioner keeps the library
// forms (its parity bounds are set against FP64 at FP32 accuracy).
template <bool FL>
__device__ __forceinline__ float fln(float x) {
  if constexpr (FL) return __builtin_amdgcn_logf(x) * 0.6931471805599453f;
  else return __logf(x);
}
template <bool FL>
__device__ __forceinline__ float fsqrt(float x) {
  if constexpr (FL) return __builtin_amdgcn_sqrtf(x);
  else return sqrtf(x);
}

// (m, j) walk over an MT x n tile by the 256 threads of a workgroup, without
// a division per element: start at tid, advance by 256 elements.
struct TileWalk {
  int m, j, dm, dj, n;
  // the (row, column) walk of nt threads over a tile of n columns
  __device__ TileWalk(int tid, int n_, int nt = 256) : n(n_) {
    m = tid / n;
    j = tid - m * n;
    dm = nt / n;
    dj = nt - dm * n;
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
// panel_prefetch issues the ring's first P k-steps and panel_run consumes
// them, so that a caller can put the first loads before a barrier.
template <class PT, int CB, int P>
struct BRing {
  typename PT::BF r[P][CB];
};

// Every ring slot is loaded unconditionally (k and column indices clamped
// to valid blocks; the MFMAs of steps past kb1 and of a missing second
// column block are skipped instead), so each group issues the same loads and
// the compiler's wait before a step covers that step's fragments only. A
// conditional load made it wait for every outstanding load (and store) at the
// top of each group.
template <class PT, int CB, int P>
__device__ __forceinline__ void panel_prefetch(BRing<PT, CB, P>& ring, const typename PT::T* Bp, int KBtot, int nb0,
                                               int nbstep, int ncb, int kb0, int kb1, int lane) {
  // (an empty range loads block kb0, or the last block: any valid block)
  const int klast = kb1 > kb0 ? kb1 - 1 : (kb0 < KBtot ? kb0 : KBtot - 1);
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int nb = j < ncb ? nb0 + j * nbstep : nb0;
      ring.r[p][j] = PT::load_b(Bp, KBtot, nb, kb0 + p < klast ? kb0 + p : klast, lane);
    }
}

template <class PT, int MB, int CB, int P, bool TR = false>
__device__ __forceinline__ void panel_run(BRing<PT, CB, P>& ring, const typename PT::T* As, int lda,
                                          const typename PT::T* Bp, int KBtot, int nb0, int nbstep, int ncb, int kb0,
                                          int kb1, int a_kb0, f32x16_t (&acc)[CB][MB], int lane) {
  using AF = decltype(PT::load_a(As, 0, 0, 0, 0));
  const int klast = kb1 - 1;
  for (int kb = kb0; kb < kb1; kb += P) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int k = kb + p;
      AF a[MB];
#pragma unroll
      for (int r = 0; r < MB; ++r) a[r] = PT::load_a(As, lda, r * 32, (k < klast ? k : klast) - kb0 + a_kb0, lane);
      if (k < kb1) {
#pragma unroll
        for (int j = 0; j < CB; ++j)
          if (j < ncb) {
#pragma unroll
            for (int r = 0; r < MB; ++r)
              acc[j][r] = TR ? PT::mma(ring.r[p][j], a[r], acc[j][r]) : PT::mma(a[r], ring.r[p][j], acc[j][r]);
          }
      }
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        const int nb = j < ncb ? nb0 + j * nbstep : nb0;
        ring.r[p][j] = PT::load_b(Bp, KBtot, nb, k + P < klast ? k + P : klast, lane);
      }
    }
  }
}

// The guarded form (loads only inside the range): for kernels that run
// several waves per SIMD, where the other waves hide the group-top wait and
// the clamped form's extra loads of short k-ranges would cost L1 bandwidth
// (the flow pass kernel, three waves per SIMD).
template <class PT, int MB, int CB, int P, bool TR = false>
__device__ __forceinline__ void mfma_panel(const typename PT::T* As, int lda, const typename PT::T* Bp, int KBtot,
                                           int nb0, int nbstep, int ncb, int kb0, int kb1, int a_kb0,
                                           f32x16_t (&acc)[CB][MB], int lane) {
  using AF = decltype(PT::load_a(As, 0, 0, 0, 0));
  typename PT::BF ring[P][CB];
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
// the adjoint runs about one wave per SIMD at training batch sizes: a deeper
// ring in the registers its packed tanh derivatives free
constexpr int BWD_PF = 4;
constexpr int BWD_PF_R = 4;
// the pass kernel fits an 8-deep ring in its three-waves-per-SIMD budget (167 VGPRs)
constexpr int PASS_PF = 8;

// arr[i] for a uniform i, by compare-and-select over the static indices (a
// dynamic index into a by-value kernel argument would copy it to scratch)
template <class V, int N>
__device__ __forceinline__ V pick(const V (&arr)[N], int i) {
  V v = arr[0];
#pragma unroll
  for (int k = 1; k < N; ++k) v = (i == k) ? arr[k] : v;
  return v;
}

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
static __device__ void spline_fwd(const float* pp, const FlowDev& f, float x, float& y, float& ld) {
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
static __device__ void spline_inv(const float* pp, const FlowDev& f, float yv, float& x, float& ld) {
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
static __device__ float spline_vjp(const float* pp, const FlowDev& f, float x, float gy, float gl, T* go, CV cvt) {
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

template <bool FL = false>
__device__ __forceinline__ float rs_softplus(float v) { return fmaxf(v, 0.f) + fln<FL>(1.f + __expf(-fabsf(v))); }

// knots of the bin holding t (t along x when INV is false, along y otherwise)
// KT > 0: the bin count is the compile-time constant KT (loops unrolled to
// exactly KT, no predication); KT == 0: runtime K <= 10, loops predicated.
// FL (the BF16 conditioner): the softmax exponentials as exp2 of one fma,
// p log2(e) - max log2(e), instead of a subtraction, a multiply and exp2.
template <bool INV, int KT, bool FL = false>
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
  constexpr float L2E = 1.4426950408889634f;
  const float mwl = mw * L2E, mhl = mh * L2E;
  float ew[10], eh[10], sw = 0.f, sh = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (FL) {
      ew[i] = (KT > 0 || i < K) ? __builtin_amdgcn_exp2f(fmaf(pp[i], L2E, -mwl)) : 0.f;
      eh[i] = (KT > 0 || i < K) ? __builtin_amdgcn_exp2f(fmaf(pp[10 + i], L2E, -mhl)) : 0.f;
    } else {
      ew[i] = (KT > 0 || i < K) ? __expf(pp[i] - mw) : 0.f;
      eh[i] = (KT > 0 || i < K) ? __expf(pp[10 + i] - mh) : 0.f;
    }
    sw += ew[i];
    sh += eh[i];
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float cw = ndiv(span, sw), ch = ndiv(span, sh);
  // right knots of bins 0..K-1 (the last one is hi)
  float xn[10], yn[10];
  float xc = f.lo, yc = f.lo;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    xn[i] = (i == K - 1) ? f.hi : xc + (ew[i] * cw + f.mb);
    yn[i] = (i == K - 1) ? f.hi : yc + (eh[i] * ch + f.mb);
    xc = xn[i];
    yc = yn[i];
  }
  // The knots increase (every bin is at least mb wide), so the bins whose
  // right knot t has reached form a prefix: start in bin 0 and move one bin
  // right, by selects, for every interior knot with t >= knot. This is the
  // first bin whose right knot exceeds t (t < hi on entry).
  RsBin b{f.lo, xn[0], f.lo, yn[0], pp[20], (K == 1) ? pp[30] : pp[21]};
#pragma unroll
  for (int i = 0; i + 1 < N; ++i) {
    if (KT > 0 || i + 1 < K) {
      const bool c = INV ? (t >= yn[i]) : (t >= xn[i]);
      b.xk = c ? xn[i] : b.xk;
      b.yk = c ? yn[i] : b.yk;
      b.uk = c ? pp[21 + i] : b.uk;
      b.xk1 = c ? xn[i + 1] : b.xk1;
      b.yk1 = c ? yn[i + 1] : b.yk1;
      b.uk1 = c ? ((i + 2 == K) ? pp[30] : pp[22 + i]) : b.uk1;
    }
  }
  return b;
}

template <int KT, bool FL = false>
__device__ __forceinline__ void rs_fwd(const float (&pp)[32], int K, const FlowDev& f, float x, float& y, float& ld) {
  if (x <= f.lo || x >= f.hi) {
    const float u = (x <= f.lo) ? pp[20] : pp[30];
    const float e = (x <= f.lo) ? f.lo : f.hi;
    const float dd = rs_softplus<FL>(u + f.off) + f.ms;
    y = (x - e) * dd + e;
    ld = fln<FL>(dd);
    return;
  }
  const RsBin b = rs_bin<false, KT, FL>(pp, K, f, x);
  const float dk = rs_softplus<FL>(b.uk + f.off) + f.ms, dk1 = rs_softplus<FL>(b.uk1 + f.off) + f.ms;
  const float w = b.xk1 - b.xk, hh = b.yk1 - b.yk, s = ndiv(hh, w);
  const float z = fminf(fmaxf(ndiv(x - b.xk, w), 0.f), 1.f);
  const float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  const float st = dk1 + dk - 2.f * s;
  const float num = hh * (s * z2 + dk * zz);
  const float den = s + st * zz;
  y = b.yk + ndiv(num, den);
  ld = 2.f * fln<FL>(s) + fln<FL>(dk1 * z2 + 2.f * s * zz + dk * q) - 2.f * fln<FL>(den);
}

template <int KT, bool FL = false>
__device__ __forceinline__ void rs_inv(const float (&pp)[32], int K, const FlowDev& f, float yv, float& x, float& ld) {
  if (yv <= f.lo || yv >= f.hi) {
    const float u = (yv <= f.lo) ? pp[20] : pp[30];
    const float e = (yv <= f.lo) ? f.lo : f.hi;
    const float dd = rs_softplus<FL>(u + f.off) + f.ms;
    x = (yv - e) / dd + e;
    ld = -fln<FL>(dd);
    return;
  }
  const RsBin b = rs_bin<true, KT, FL>(pp, K, f, yv);
  const float dk = rs_softplus<FL>(b.uk + f.off) + f.ms, dk1 = rs_softplus<FL>(b.uk1 + f.off) + f.ms;
  const float w = b.xk1 - b.xk, hh = b.yk1 - b.yk, s = ndiv(hh, w);
  const float wv = fminf(fmaxf(ndiv(yv - b.yk, hh), 0.f), 1.f);
  const float st = dk1 + dk - 2.f * s;
  const float c = -s * wv;
  const float bq = dk - st * wv;
  const float a = s - bq;
  const float disc = bq * bq - 4.f * a * c;
  const float sq = disc > 0.f ? fsqrt<FL>(fmaxf(disc, 1.17549435e-38f)) : 0.f;
  // the cancellation-free root: 2c / (-b - sqrt) for b >= 0, else
  // (-b + sqrt) / 2a; operands selected first, one division for both
  const bool bpos = bq >= 0.f;
  float z = ndiv(bpos ? 2.f * c : -bq + sq, bpos ? -bq - sq : 2.f * a);
  z = fminf(fmaxf(z, 0.f), 1.f);
  x = w * z + b.xk;
  const float z2 = z * z, zz = z - z2, q = (1.f - z) * (1.f - z);
  const float den = s + st * zz;
  ld = -2.f * fln<FL>(s) - fln<FL>(dk1 * z2 + 2.f * s * zz + dk * q) + 2.f * fln<FL>(den);
}

// Register-spline forward + vector-Jacobian product (the reverse sweep of
// spline_vjp): given dL/dy (gy) and dL/dlogdet (gl), writes dL/d(slot
// logits) to go (zero in padding slots) and returns dL/dx.
template <int KT, bool FL = false>
__device__ __forceinline__ float rs_vjp(const float (&pp)[32], int Kr, const FlowDev& f, float x, float gy, float gl,
                                        float (&go)[32]) {
  constexpr int N = KT > 0 ? KT : 10;
  const int K = KT > 0 ? KT : Kr;
#pragma unroll
  for (int i = 0; i < 32; ++i) go[i] = 0.f;
  if (x <= f.lo || x >= f.hi) {
    const bool left = x <= f.lo;
    const float u = (left ? pp[20] : pp[30]) + f.off;
    const float dd = rs_softplus<FL>(u) + f.ms;
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
  constexpr float L2E = 1.4426950408889634f;
  const float mwl = mw * L2E, mhl = mh * L2E;
  float ew[10], eh[10], sw = 0.f, sh = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (FL) {  // exp2 of one fma, as rs_bin
      ew[i] = (KT > 0 || i < K) ? __builtin_amdgcn_exp2f(fmaf(pp[i], L2E, -mwl)) : 0.f;
      eh[i] = (KT > 0 || i < K) ? __builtin_amdgcn_exp2f(fmaf(pp[10 + i], L2E, -mhl)) : 0.f;
    } else {
      ew[i] = (KT > 0 || i < K) ? __expf(pp[i] - mw) : 0.f;
      eh[i] = (KT > 0 || i < K) ? __expf(pp[10 + i] - mh) : 0.f;
    }
    sw += ew[i];
    sh += eh[i];
  }
  const float span = (f.hi - f.lo) - K * f.mb;
  const float isw = ndiv(1.f, sw), ish = ndiv(1.f, sh);
  const float cw = span * isw, ch = span * ish;
  // softmax masses and right knots of bins 0..K-1
  float pw[10], ph[10], xn[10], yn[10];
  float xc = f.lo, yc = f.lo;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    pw[i] = ew[i] * isw;
    ph[i] = eh[i] * ish;
    xn[i] = (i == K - 1) ? f.hi : xc + (ew[i] * cw + f.mb);
    yn[i] = (i == K - 1) ? f.hi : yc + (eh[i] * ch + f.mb);
    xc = xn[i];
    yc = yn[i];
  }
  // bin walk by selects over the increasing interior knots, as rs_bin; the
  // masses of the bins passed (< k) are summed in bin order
  float xk = f.lo, yk = f.lo, xk1 = xn[0], yk1 = yn[0], uk = pp[20], uk1 = (K == 1) ? pp[30] : pp[21];
  float pw_lt = 0.f, ph_lt = 0.f, pwk = pw[0], phk = ph[0];  // softmax mass of bins < k, and of bin k
  int k = 0;
#pragma unroll
  for (int i = 0; i + 1 < N; ++i) {
    if (KT > 0 || i + 1 < K) {
      const bool c = x >= xn[i];
      k = c ? i + 1 : k;
      xk = c ? xn[i] : xk;
      yk = c ? yn[i] : yk;
      uk = c ? pp[21 + i] : uk;
      xk1 = c ? xn[i + 1] : xk1;
      yk1 = c ? yn[i + 1] : yk1;
      uk1 = c ? ((i + 2 == K) ? pp[30] : pp[22 + i]) : uk1;
      pw_lt = c ? pw_lt + pw[i] : pw_lt;
      ph_lt = c ? ph_lt + ph[i] : ph_lt;
      pwk = c ? pw[i + 1] : pwk;
      phk = c ? ph[i + 1] : phk;
    }
  }
  uk += f.off;
  uk1 += f.off;
  const float dk = rs_softplus<FL>(uk) + f.ms, dk1 = rs_softplus<FL>(uk1) + f.ms;
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
      go[m] = pw[m] * (gwm - dotw);
      go[10 + m] = ph[m] * (ghm - doth);
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
  int xs, ldr, ldp, r1, r2, ub, total;
  int ldx, lh, lo_, la0;
  __host__ __device__ PassSmem(const FlowDims& dm, int MT) {
    ldx = dm.d + 1;
    lh = dm.hmaxp() + PT::PAD;  // hidden activations (either region)
    lo_ = dm.ncf() + 1;
    la0 = rup(dm.kc(1) > dm.kc(0) ? dm.kc(1) : dm.kc(0), PT::KSTEP) + PT::PAD;
    int o = 0;
    xs = o;
    o += rup(MT * ldx * 4, 16);
    ldr = o;
    o += rup(MT * 4, 16);
    ldp = o;  // per-dim log-dets of a chunk, or the (up to 8) waves' partial log-dets
    o += rup(MT * (dm.G() > 8 ? dm.G() : 8) * 4, 16);
    // two regions, each holding one of: hidden activations, the conditioner
    // input (region 2 at the start of a layer) or a chunk's spline
    // parameters (the region that does not hold the last hidden layer)
    int sh = MT * lh * (int)sizeof(typename PT::T), sa = MT * la0 * (int)sizeof(typename PT::T);
    int so = dm.reg_spline() ? 0 : MT * lo_ * 4;
    int sr = rup(sh > sa ? (sh > so ? sh : so) : (sa > so ? sa : so), 16);
    r1 = o;
    o += sr;
    r2 = o;
    o += sr;
    ub = o;  // dense base Gaussian: base noise / whitened residuals of the tile [MT][ldx] (FP32)
    o += dm.base_dense ? rup(MT * ldx * 4, 16) : 0;
    total = o;
  }
};

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
  int ldx, lh, lo_, lg;
  __host__ __device__ BwdSmem(const FlowDims& dm, int MT) {
    ldx = dm.d + 1;
    lh = dm.hmaxp() + PT::PAD;  // hidden activations / their gradients (two regions)
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
    // with the register spline, the hidden buffer that does not hold the
    // last hidden layer also stages every other round's parameter gradients
    const int hbb = rup(MT * (dm.reg_spline() && lg > lh ? lg : lh) * tb, 16);
    hb1 = o;
    o += hbb;
    hb2 = o;
    o += hbb;
    ob = o;  // spline parameters of a chunk (chunked path only)
    o += dm.reg_spline() ? 0 : rup(MT * lo_ * 4, 16);
    gpb = o;
    o += rup(MT * lg * tb, 16);
    red = o;
    o += 64;
    total = o;
  }
};

// compile-time bin count for the register spline: 8 and 10 (the common
// configurations), 0 = runtime K (any K <= 10, and the chunked path)
inline int spline_kt(int K) { return K == 8 ? 8 : (K == 10 ? 10 : 0); }

// --------------------------------------------------------------- host ----
// compute units of the current device (queried once; 256 if the query fails)
static inline int device_cus() {
  static int n = 0;
  if (n == 0) {
    int dev = 0, cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0)
      n = cu;
    else
      n = 256;
  }
  return n;
}

static int check_flow(const fmc_flow* fl, FlowDev& fd) {
  FMC_ARG(fl != nullptr, "flow is NULL");
  FMC_ARG(fl->n_dim >= 2 && fl->n_dim <= 512, "flow n_dim must be in [2, 512] (got %d)", fl->n_dim);
  FMC_ARG(fl->n_layers >= 1 && fl->n_layers <= 256, "flow n_layers must be in [1, 256]");
  FMC_ARG(fl->n_bins >= 1 && fl->n_bins <= 42, "flow n_bins must be in [1, 42]");
  FMC_ARG(fl->n_hidden >= 1 && fl->n_hidden <= FLOW_MAX_HIDDEN,
          "conditioners need 1 to %d hidden layers (got %d)", FLOW_MAX_HIDDEN, fl->n_hidden);
  for (int i = 0; i < fl->n_hidden; ++i)
    FMC_ARG(fl->hidden[i] >= 1 && fl->hidden[i] <= 512, "hidden size %d must be in [1, 512] (got %d)", i,
            fl->hidden[i]);
  FMC_ARG(fl->precision == FMC_PREC_F32 || fl->precision == FMC_PREC_BF16, "unknown precision %d", fl->precision);
  FMC_ARG(fl->theta != nullptr, "flow theta is NULL");
  FMC_ARG(fl->range_max > fl->range_min, "spline range must be increasing");
  FMC_ARG(fl->n_bins * fl->min_bin_size < fl->range_max - fl->range_min, "min_bin_size too large for the range");
  FMC_ARG(fl->min_knot_slope > 0.f && fl->min_knot_slope < 1.f, "min_knot_slope must be in (0, 1)");
  FlowDims dm{};
  dm.d = fl->n_dim;
  dm.L = fl->n_layers;
  dm.K = fl->n_bins;
  dm.nh = fl->n_hidden;
  for (int i = 0; i < FLOW_MAX_HIDDEN; ++i) dm.h[i] = i < fl->n_hidden ? fl->hidden[i] : 0;
  dm.kstep = fl->precision == FMC_PREC_BF16 ? 16 : 2;
  FMC_ARG((fl->flags & ~FMC_FLOW_DENSE_BASE) == 0, "unknown flow flags 0x%x", fl->flags);
  dm.base_dense = (fl->flags & FMC_FLOW_DENSE_BASE) ? 1 : 0;
  dm.finalize();
  fd.dm = dm;
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
