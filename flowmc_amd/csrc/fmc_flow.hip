// MaskedCouplingRQSpline flow on gfx950: packed MFMA conditioner GEMMs with
// the rational-quadratic spline fused as the epilogue of the last GEMM.
//
//   flow_pass_kernel      (fmc_flow_pass.hip) all layers of a forward /
//                         inverse pass for a tile of 32*MB samples
//   flow_bwd_kernel_*     (fmc_flow_bwd.hip) adjoint of one forward layer
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
#include "fmc_flow_common.h"

#include <type_traits>

namespace fmc {

// --------------------------------------------------- weight gradients ----
// One workgroup of WG_SPLIT wavefronts per 32x32 tile of gW = G^T Act: the
// batch reduction is split over the wavefronts (each takes an interleaved
// share of the KSTEP-wide batch slices) and the partial tiles are summed in
// LDS in a fixed order. Both operands come from the fragment-blocked store:
// each fragment is one coalesced 64-lane read.
// Problems: i = 0 .. nh-1 the hidden layers' W_i (W_0 restricted to the
// conditioning columns), nh the output layer W_nh (gradient-store order).
// Tiles of the first column block (kb == 0) also sum their G rows into the
// bias gradients. The last tile sums the ScalarAffine partials. Workgroup b
// runs on XCD b % 8; tiles are dealt so that XCD x gets a contiguous range
// of them, i.e. the tiles sharing a G row panel share one L2.
constexpr int WG_SPLIT = 8;
static_assert(16 % WG_SPLIT == 0, "each wave finishes a whole number of accumulator registers");
constexpr int WG_XCDS = 8;
constexpr int WG_PF = 8;  // weight-gradient prefetch depth (batch slices)

// Problem table of one layer, built on the host: problem i covers tiles
// start[i] .. start[i+1]-1, its G rows (gradient of the problem's output)
// begin at store row g_row[i] and its A rows (its input) at a_row[i], with
// n_a[i] / 32 column blocks.
constexpr int WG_MAXP = FLOW_MAX_HIDDEN + 1;
struct WgTable {
  int total;  // tiles, including the ScalarAffine tile
  int start[WG_MAXP + 1], g_row[WG_MAXP], a_row[WG_MAXP], n_a[WG_MAXP];
  // theta offsets of the problem's weight and bias, its output rows and its
  // weight row stride (input width)
  int64_t th_w[WG_MAXP], th_b[WG_MAXP];
  int n_out[WG_MAXP], in_w[WG_MAXP];
};


static WgTable wgrad_table(const FlowDims& dm, int l) {
  WgTable t{};
  int s = 0;
  for (int i = 0; i <= dm.nh; ++i) {
    int g_row, n_g, a_row, n_a;
    if (i == dm.nh) {
      g_row = (int)dm.st_gp(l), n_g = dm.n3(l), a_row = (int)dm.st_h(l, dm.nh - 1), n_a = dm.hlp();
    } else if (i == 0) {
      g_row = (int)dm.st_gh(l, 0), n_g = dm.hp(0), a_row = (int)dm.st_c(l), n_a = dm.ncp(l);
    } else {
      g_row = (int)dm.st_gh(l, i), n_g = dm.hp(i), a_row = (int)dm.st_h(l, i - 1), n_a = dm.hp(i - 1);
    }
    t.start[i] = s;
    t.g_row[i] = g_row;
    t.a_row[i] = a_row;
    t.n_a[i] = n_a;
    t.th_w[i] = dm.th_w(l, i);
    t.th_b[i] = dm.th_b(l, i);
    t.n_out[i] = i == dm.nh ? dm.d * dm.P() : dm.h[i];
    t.in_w[i] = dm.lin_in(i);
    s += (n_g / 32) * (n_a / 32);
  }
  for (int i = dm.nh + 1; i <= WG_MAXP; ++i) t.start[i] = s;
  t.total = s + 1;
  return t;
}

template <class PT>
__global__ void __launch_bounds__(64 * WG_SPLIT) flow_wgrad_kernel(FlowDev f, int l, WgTable tab, const void* store,
                                                                   int64_t ld, const float* aff_part, int n_part,
                                                                   float* grad) {
  using T = typename PT::T;
  __shared__ float red[WG_SPLIT][16][64];
  __shared__ float rsum[WG_SPLIT][64];
  const FlowDims& dm = f.dm;
  // the wave index in a scalar register: branches on it are uniform
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int total = tab.total;
  const int per = (total + WG_XCDS - 1) / WG_XCDS;
  int tt = (int)(blockIdx.x % WG_XCDS) * per + (int)(blockIdx.x / WG_XCDS);
  if (tt >= total) return;
  if (tt == total - 1) {  // ScalarAffine gradients: fixed-order sum of the partials
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
  int prob = 0;
#pragma unroll
  for (int i = 1; i < WG_MAXP; ++i) prob += (tt >= tab.start[i]) ? 1 : 0;
  tt -= pick(tab.start, prob);
  const int gRow = pick(tab.g_row, prob), aRow = pick(tab.a_row, prob), ncolb = pick(tab.n_a, prob) / 32;
  const int64_t thw = pick(tab.th_w, prob), thb = pick(tab.th_b, prob);
  const int n_out = pick(tab.n_out, prob), in_w = pick(tab.in_w, prob);
  const bool out_layer = prob == dm.nh;
  const int nb = tt / ncolb, kb = tt - nb * ncolb;
  const bool want_bias = (kb == 0);
  const T* st = reinterpret_cast<const T*>(store);
  f32x16_t acc = {};
  float bsum = 0.f;
  // this wave takes batch slices wave, wave + WG_SPLIT, ... ; fragments
  // (L2-resident store) are prefetched P slices ahead. Every full group of
  // P slices but the last refills its ring slots unconditionally, so the
  // wait before a slice's MFMA covers that slice's loads only.
  constexpr int P = WG_PF;
  using GF = decltype(PT::load_g(st, 0, 0, 0, 0));
  GF ga[P], ab[P];
  const int nsteps = (int)(ld / PT::KSTEP);
  const int gRb = gRow / 32 + nb, aRb = aRow / 32 + kb;
  const int my = (nsteps - wave + WG_SPLIT - 1) / WG_SPLIT;  // slices of this wave
  // the prologue loads every slot (slots past the wave's last slice repeat
  // it, or read slice 0 when the wave has none), so the loop's waits stay exact
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int sl = my > 0 ? wave + (p < my ? p : my - 1) * WG_SPLIT : 0;
    ga[p] = PT::load_g(st, nsteps, gRb, sl, lane);
    ab[p] = PT::load_g(st, nsteps, aRb, sl, lane);
  }
  const int ng = my / P, tail = my - ng * P;
  int i = 0;
  for (int g = 0; g + 1 < ng; ++g) {
#pragma unroll
    for (int p = 0; p < P; ++p, ++i) {
      acc = PT::mma(ga[p], ab[p], acc);
      if (want_bias) bsum += PT::hsum(ga[p]);
      const int sl = wave + (i + P) * WG_SPLIT;
      ga[p] = PT::load_g(st, nsteps, gRb, sl, lane);
      ab[p] = PT::load_g(st, nsteps, aRb, sl, lane);
    }
  }
  if (ng > 0) {  // the last full group refills the tail's slots only
#pragma unroll
    for (int p = 0; p < P; ++p, ++i) {
      acc = PT::mma(ga[p], ab[p], acc);
      if (want_bias) bsum += PT::hsum(ga[p]);
      if (p < tail) {
        const int sl = wave + (i + P) * WG_SPLIT;
        ga[p] = PT::load_g(st, nsteps, gRb, sl, lane);
        ab[p] = PT::load_g(st, nsteps, aRb, sl, lane);
      }
    }
  }
  const int nt = ng > 0 ? tail : my;  // slices left, in ring slots 0 .. nt-1
#pragma unroll
  for (int p = 0; p < P - 1; ++p)
    if (p < nt) {
      acc = PT::mma(ga[p], ab[p], acc);
      if (want_bias) bsum += PT::hsum(ga[p]);
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
    if (out_layer) {
      int row = dm.gp_row(l, n);
      if (row >= 0 && k < in_w) idx = thw + (int64_t)row * in_w + k;
    } else if (prob == 0) {
      if (n < n_out && k < dm.dc(l)) idx = thw + (int64_t)n * in_w + dm.cond(l, k);
    } else {
      if (n < n_out && k < in_w) idx = thw + (int64_t)n * in_w + k;
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
      if (out_layer) {
        int row = dm.gp_row(l, n);
        if (row >= 0) idx = thb + row;
      } else if (n < n_out) {
        idx = thb + n;
      }
      if (idx >= 0) grad[idx] = b;
    }
  }
}

// ----------------------------------------- split weight gradients ----
// gW = G^T Act of one layer as a split-K product: the batch (the reduction
// axis, 10^4 samples) is cut into S ranges and every (block, range) pair is
// one workgroup of 8 waves. A block is WS_N G row blocks x up to WS_K A row
// blocks (128 x 256 elements of gW); per batch slice the block's 12 store
// fragments are staged in LDS once (double-buffered, WS_KS slices a stage:
// 48 KB a workgroup, which fits beside an adjoint workgroup's 68 KB on a CU)
// and wave w multiplies G block w/2 with A blocks 4(w%2) .. +3, so a 1 KB
// fragment read from L2 feeds 2.7 MFMAs instead of 0.5. The FP32 partial
// blocks go to a workspace and flow_wgrad_reduce_kernel sums the S partials
// of each element in a fixed order (deterministic, like the one-workgroup-
// per-tile kernel above, whose sums it matches to rounding).
constexpr int WS_N = 4, WS_K = 8, WS_KS = 2, WS_FR = WS_N + WS_K, WS_WAVES = 8;
constexpr int WS_PART = WS_N * 32 * WS_K * 32 + WS_N * 32;  // floats of one partial (block + bias rows)

struct WsTable {
  WgTable wt;                 // the layer's problems
  int blk_start[WG_MAXP + 1];  // first block of problem i
  int nbk[WG_MAXP];           // k-chunks of problem i (blocks = n-chunks x k-chunks)
  int n_blocks, S;
};

static WsTable wsplit_table(const FlowDims& dm, int l, int64_t ld, int max_wg) {
  WsTable t{};
  t.wt = wgrad_table(dm, l);
  int b = 0;
  for (int i = 0; i <= dm.nh; ++i) {
    const int n_g = (t.wt.start[i + 1] - t.wt.start[i]) / (t.wt.n_a[i] / 32) * 32;  // G rows of problem i
    const int nbn = (n_g / 32 + WS_N - 1) / WS_N, nbk = (t.wt.n_a[i] / 32 + WS_K - 1) / WS_K;
    t.blk_start[i] = b;
    t.nbk[i] = nbk;
    b += nbn * nbk;
  }
  for (int i = dm.nh + 1; i <= WG_MAXP; ++i) t.blk_start[i] = b;
  t.n_blocks = b;
  const int ns = (int)(ld / 16);  // batch slices of a BF16 store (FP32: 8x as many; the split stays)
  // S a multiple of the XCD count where the grid allows: workgroup b runs on
  // XCD b % 8 and the workgroups of batch range s (all blocks) then share
  // XCD s % 8, so the A-row slices every block of that range reads come
  // from one L2 (a range spread over eight XCDs was read from memory by each)
  int S = max_wg / b;
  if (S >= WG_XCDS) S -= S % WG_XCDS;
  S = S < 1 ? 1 : S;
  const int smax = ns / (2 * WS_KS);  // at least two stages per range
  S = S > smax ? (smax < 1 ? 1 : smax) : S;
  t.S = S;
  return t;
}

template <class PT>
__global__ void __launch_bounds__(64 * WS_WAVES) flow_wgrad_split_kernel(FlowDev f, WsTable tab, const void* store,
                                                                         int64_t ld, float* part) {
  using T = typename PT::T;
  using GF = decltype(PT::load_g(static_cast<const T*>(nullptr), 0, 0, 0, 0));
  constexpr int FB = 64 * (int)sizeof(GF);  // bytes of one fragment
  extern __shared__ __align__(16) unsigned char wsm[];
  const FlowDims& dm = f.dm;
  // the wave index in a scalar register: branches on it are uniform
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int S = tab.S;
  const int blk = blockIdx.x / S, sidx = blockIdx.x - blk * S;
  int prob = 0;
#pragma unroll
  for (int i = 1; i < WG_MAXP; ++i) prob += (blk >= tab.blk_start[i]) ? 1 : 0;
  const int nbk = pick(tab.nbk, prob);
  const int lb = blk - pick(tab.blk_start, prob);
  const int nch = lb / nbk, kch = lb - nch * nbk;
  const int n_a = pick(tab.wt.n_a, prob);
  const int n_g = (pick(tab.wt.start, prob + 1) - pick(tab.wt.start, prob)) / (n_a / 32) * 32;
  const int nbs = (n_g / 32 - nch * WS_N) < WS_N ? (n_g / 32 - nch * WS_N) : WS_N;  // G row blocks of the block
  const int kbs = (n_a / 32 - kch * WS_K) < WS_K ? (n_a / 32 - kch * WS_K) : WS_K;  // A row blocks
  const int gRb0 = pick(tab.wt.g_row, prob) / 32 + nch * WS_N, aRb0 = pick(tab.wt.a_row, prob) / 32 + kch * WS_K;
  const T* st = reinterpret_cast<const T*>(store);
  const int ns = (int)(ld / PT::KSTEP);
  const int s0 = (int)((int64_t)sidx * ns / S), s1 = (int)((int64_t)(sidx + 1) * ns / S);
  const int nstage = (s1 - s0 + WS_KS - 1) / WS_KS;
  // wave roles: G block wn, A blocks wk0 .. wk0+3
  const int wn = wave >> 1, wk0 = (wave & 1) * 4;
  const bool want_bias = (wave & 1) == 0 && kch == 0;
  // this wave's staged fragments: f = wave + 8 u (u < 6) of the stage's 48
  constexpr int NU = WS_KS * WS_FR / WS_WAVES;
  static_assert(WS_KS * WS_FR % WS_WAVES == 0, "stage fragments split evenly over the waves");
  // two register sets of staged fragments: the loads of stage s + 2 are
  // issued before stage s is multiplied and written to LDS after stage s + 1
  // is, so each stage's loads have two stages' MFMAs to land. Every load is
  // unconditional (stages past the range repeat its last slice), so the
  // wait before a stage's LDS write covers that stage's loads only.
  GF rg[2][NU];
  auto load_stage = [&](GF (&dst)[NU], int stg) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int fr = wave + WS_WAVES * u, j = fr / WS_FR, which = fr - j * WS_FR;
      int sl = s0 + stg * WS_KS + j;
      sl = sl < s1 ? sl : s1 - 1;
      const int rb = which < WS_N ? gRb0 + (which < nbs ? which : nbs - 1)
                                  : aRb0 + (which - WS_N < kbs ? which - WS_N : kbs - 1);
      dst[u] = PT::load_g(st, ns, rb, sl, lane);
    }
  };
  auto put_stage = [&](const GF (&src)[NU], int buf) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int fr = wave + WS_WAVES * u;
      *reinterpret_cast<GF*>(wsm + ((size_t)(buf * WS_KS * WS_FR + fr) * 64 + lane) * sizeof(GF)) = src[u];
    }
  };
  auto frag = [&](int buf, int j, int which) -> GF {
    return *reinterpret_cast<const GF*>(wsm + ((size_t)((buf * WS_KS + j) * WS_FR + which) * 64 + lane) * sizeof(GF));
  };
  f32x16_t acc[4] = {};
  float bsum = 0.f;
  load_stage(rg[0], 0);
  load_stage(rg[1], 1);
  put_stage(rg[0], 0);
  __syncthreads();
  // one stage on LDS buffer / register set CUR (a template argument, so the
  // register sets are indexed statically)
  auto stage = [&](auto cur_c, int stg) {
    constexpr int cur = decltype(cur_c)::value;
    load_stage(rg[cur], stg + 2);  // (rg[cur] was written to LDS at the previous stage)
#pragma unroll
    for (int j = 0; j < WS_KS; ++j) {
      if (s0 + stg * WS_KS + j < s1 && wn < nbs) {
        const GF g = frag(cur, j, wn);
        if (want_bias) bsum += PT::hsum(g);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          if (wk0 + kk < kbs) acc[kk] = PT::mma(g, frag(cur, j, WS_N + wk0 + kk), acc[kk]);
      }
    }
    put_stage(rg[cur ^ 1], cur ^ 1);
    __syncthreads();
  };
  // an even number of stages (a padding stage multiplies nothing)
  for (int stg = 0; stg < nstage; stg += 2) {
    stage(std::integral_constant<int, 0>{}, stg);
    stage(std::integral_constant<int, 1>{}, stg + 1);
  }
  // partial block [n_local < 128][k_local < 256] and bias rows [128]
  float* pp = part + (size_t)blockIdx.x * WS_PART;
  if (wn < nbs) {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = (wk0 + kk) * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) pp[(wn * 32 + out_row(v, lane)) * (WS_K * 32) + k] = acc[kk][v];
    }
    if (!(wave & 1)) {
      // lane l and l+32 hold the two halves of G row wn*32 + l%32
      const float b = bsum + __shfl_xor(bsum, 32, 64);
      if (lane < 32) pp[WS_N * 32 * WS_K * 32 + wn * 32 + lane] = b;
    }
  }
}

// Fixed-order sum of the S partials of every element of every block, mapped
// to theta (as the epilogue of flow_wgrad_kernel), plus the ScalarAffine sums.
__global__ void __launch_bounds__(256) flow_wgrad_reduce_kernel(FlowDev f, int l, WsTable tab, const float* part,
                                                                const float* aff_part, int n_part, float* grad) {
  const FlowDims& dm = f.dm;
  if (blockIdx.x == gridDim.x - 1) {
    // ScalarAffine gradients: thread t sums a contiguous run of the
    // partials, then a fixed pairwise tree over the 256 runs
    __shared__ float r0[256], r1[256];
    const int t = threadIdx.x, chunk = (n_part + 255) / 256;
    float s0 = 0.f, s1 = 0.f;
    for (int i = t * chunk; i < (t + 1) * chunk && i < n_part; ++i) {
      s0 += aff_part[2 * i];
      s1 += aff_part[2 * i + 1];
    }
    r0[t] = s0;
    r1[t] = s1;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) {
        r0[t] += r0[t + h];
        r1[t] += r1[t + h];
      }
      __syncthreads();
    }
    if (t == 0) {
      grad[dm.th_scale(l)] = r0[0];
      grad[dm.th_shift(l)] = r1[0];
    }
    return;
  }
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int blk = (int)(e / WS_PART), q = (int)(e - (int64_t)blk * WS_PART);
  if (blk >= tab.n_blocks) return;
  int prob = 0;
#pragma unroll
  for (int i = 1; i < WG_MAXP; ++i) prob += (blk >= tab.blk_start[i]) ? 1 : 0;
  const int nbk = pick(tab.nbk, prob);
  const int lb = blk - pick(tab.blk_start, prob);
  const int nch = lb / nbk, kch = lb - nch * nbk;
  const int n_a = pick(tab.wt.n_a, prob);
  const int n_g = (pick(tab.wt.start, prob + 1) - pick(tab.wt.start, prob)) / (n_a / 32) * 32;
  const bool bias = q >= WS_N * 32 * WS_K * 32;
  const int nl = bias ? q - WS_N * 32 * WS_K * 32 : q / (WS_K * 32), kl = bias ? 0 : q - nl * (WS_K * 32);
  const int n = nch * WS_N * 32 + nl, k = kch * WS_K * 32 + kl;
  if (n >= n_g || k >= n_a || (bias && kch != 0)) return;
  const int S = tab.S;
  const float* pp = part + (size_t)blk * S * WS_PART + q;
  // loads of 8 partials issued together, summed in the fixed order 0 .. S-1
  float t = 0.f;
  int si = 0;
  for (; si + 8 <= S; si += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = pp[(size_t)(si + u) * WS_PART];
#pragma unroll
    for (int u = 0; u < 8; ++u) t += v[u];
  }
  for (; si < S; ++si) t += pp[(size_t)si * WS_PART];
  const int64_t thw = pick(tab.wt.th_w, prob), thb = pick(tab.wt.th_b, prob);
  const int n_out = pick(tab.wt.n_out, prob), in_w = pick(tab.wt.in_w, prob);
  const bool out_layer = prob == dm.nh;
  int64_t idx = -1;
  if (bias) {
    if (out_layer) {
      const int row = dm.gp_row(l, n);
      if (row >= 0) idx = thb + row;
    } else if (n < n_out) {
      idx = thb + n;
    }
  } else if (out_layer) {
    const int row = dm.gp_row(l, n);
    if (row >= 0 && k < in_w) idx = thw + (int64_t)row * in_w + k;
  } else if (prob == 0) {
    if (n < n_out && k < dm.dc(l)) idx = thw + (int64_t)n * in_w + dm.cond(l, k);
  } else {
    if (n < n_out && k < in_w) idx = thw + (int64_t)n * in_w + k;
  }
  if (idx >= 0) grad[idx] = t;
}

// ------------------------------------------------------------- packing ---
// Grid: x strides over the elements of one image, y = layer * NI + image with
// NI = 2 nh + 5 images per layer: wh(0..nh-1), w3, wht(0..nh-1), w3t, w3r,
// w3rt, and the layer's packed biases last, so every element decodes its own
// (block, lane, k) only. The per-image geometry comes from a host-built
// table (by layer parity: the conditioning / transformed dim counts differ).
constexpr int PK_MAXI = 2 * FLOW_MAX_HIDDEN + 4;  // weight images per layer
struct PackTable {
  int n_img;                                   // weight images (2 nh + 4)
  int kind[PK_MAXI], hid[PK_MAXI];             // 0 wh, 1 w3, 2 wht, 3 w3t, 4 w3r, 5 w3rt; hidden index
  int64_t off[2][PK_MAXI];                     // offset within the layer's packed block
  int n[2][PK_MAXI], kd[2][PK_MAXI];           // image rows N and K
  int64_t thw[FLOW_MAX_HIDDEN];                // hidden layer i: W offset within a theta layer
  int64_t thb[FLOW_MAX_HIDDEN];                // and its bias offset
  int rows[FLOW_MAX_HIDDEN], in_w[FLOW_MAX_HIDDEN];  // h[i], its input width
  int pbb[FLOW_MAX_HIDDEN + 1];                // packed-bias offsets of the hidden layers
  // layer-level sizes and offsets (index: layer parity), so the kernel
  // needs no layout arithmetic over the hidden-layer arrays
  int64_t pk_size[2], pb_size[2], pb_b3r[2];
  int64_t pb_b3, th_layer, th_w3, th_b3;       // th_w3 / th_b3 within a theta layer
};

static PackTable pack_table(const FlowDims& dm) {
  PackTable t{};
  const int nh = dm.nh;
  t.n_img = 2 * nh + 4;
  for (int w = 0; w < t.n_img; ++w) {
    int kind, i = 0;
    if (w < nh) {
      kind = 0, i = w;
    } else if (w == nh) {
      kind = 1;
    } else if (w < 2 * nh + 1) {
      kind = 2, i = w - nh - 1;
    } else {
      kind = 3 + (w - 2 * nh - 1);
    }
    t.kind[w] = kind;
    t.hid[w] = i;
    for (int par = 0; par < 2; ++par) {
      const int l = par;  // offsets within a layer depend on its parity only
      int64_t off;
      int N, Kd;
      switch (kind) {
        case 0: off = dm.pk_wh(l, i); N = dm.hp(i); Kd = dm.wh_k(l, i); break;
        case 1: off = dm.pk_w3(l); N = dm.n3p(l); Kd = dm.hlp(); break;
        case 2: off = dm.pk_wht(l, i); N = dm.wht_n(l, i); Kd = dm.hp(i); break;
        case 3: off = dm.pk_w3t(l); N = dm.hlp(); Kd = dm.n3p(l); break;
        case 4: off = dm.pk_w3r(l); N = dm.n3r(l); Kd = dm.hlp(); break;
        default: off = dm.pk_w3rt(l); N = dm.hlp(); Kd = dm.n3r(l); break;
      }
      t.off[par][w] = off;
      t.n[par][w] = N;
      t.kd[par][w] = Kd;
    }
  }
  for (int i = 0; i < FLOW_MAX_HIDDEN; ++i) {
    t.thw[i] = i < nh ? dm.th_lin(i) : 0;
    t.thb[i] = i < nh ? dm.th_lin(i) + (int64_t)dm.lin_out(i) * dm.lin_in(i) : 0;
    t.rows[i] = i < nh ? dm.h[i] : 0;
    t.in_w[i] = i < nh ? dm.lin_in(i) : 0;
  }
  for (int i = 0; i <= FLOW_MAX_HIDDEN; ++i) t.pbb[i] = i <= nh ? (int)dm.pb_b(i) : (int)dm.pb_b(nh);
  for (int par = 0; par < 2; ++par) {
    t.pk_size[par] = dm.pk_size(par);
    t.pb_size[par] = dm.pb_size(par);
    t.pb_b3r[par] = dm.pb_b3r(par);
  }
  t.pb_b3 = dm.pb_b3(0);
  t.th_layer = dm.th_layer_size();
  t.th_w3 = dm.th_lin(nh);
  t.th_b3 = dm.th_lin(nh) + (int64_t)dm.lin_out(nh) * dm.lin_in(nh);
  return t;
}

template <class PT>
__global__ void __launch_bounds__(256) flow_pack_kernel(FlowDev f, PackTable tab, typename PT::T* packed,
                                                        float* pbias) {
  const FlowDims& dm = f.dm;
  const int NI = tab.n_img + 1;
  const int l = blockIdx.y / NI, which = blockIdx.y - NI * l, par = l & 1;
  const int hl = dm.hl();
  const float* thl = f.theta + (int64_t)l * tab.th_layer;  // this layer's leaves
  const float* w3 = thl + tab.th_w3;
  const float* b3 = thl + tab.th_b3;
  const int64_t pk_layer = (int64_t)(l / 2) * (tab.pk_size[0] + tab.pk_size[1]) + (par ? tab.pk_size[0] : 0);
  if (which == NI - 1) {
    const int64_t n = par ? tab.pb_size[1] : tab.pb_size[0];
    float* out = pbias + (int64_t)(l / 2) * (tab.pb_size[0] + tab.pb_size[1]) + (par ? tab.pb_size[0] : 0);
    const int64_t pb3 = tab.pb_b3, pb3r = par ? tab.pb_b3r[1] : tab.pb_b3r[0];
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
      float v = 0.f;
      if (e < pb3) {
        int i = 0;
#pragma unroll
        for (int k = 1; k < FLOW_MAX_HIDDEN; ++k) i += (e >= tab.pbb[k] && k < dm.nh) ? 1 : 0;
        const int q = (int)(e - pick(tab.pbb, i));
        if (q < pick(tab.rows, i)) v = thl[pick(tab.thb, i) + q];
      } else if (e < pb3r) {
        int row = dm.w3_row(l, (int)(e - pb3));
        if (row >= 0) v = b3[row];
      } else {
        int row = dm.b3r_row(l, (int)(e - pb3r));
        if (row >= 0) v = b3[row];
      }
      out[e] = v;
    }
    return;
  }
  const int kind = pick(tab.kind, which), i = pick(tab.hid, which);
  const int64_t off = par ? pick(tab.off[1], which) : pick(tab.off[0], which);
  const int N = par ? pick(tab.n[1], which) : pick(tab.n[0], which);
  const int Kdim = par ? pick(tab.kd[1], which) : pick(tab.kd[0], which);
  // One image holds at most a few hundred thousand elements (the host checks
  // that it fits 31 bits), so the decode runs in 32-bit arithmetic. A thread
  // builds one lane's EPL consecutive elements (one MFMA B fragment, a single
  // vector store); the source row of the untransposed images (kinds 0, 1, 4)
  // is decoded once per fragment.
  static_assert((PT::EPL & (PT::EPL - 1)) == 0 && sizeof(typename PT::BF) == PT::EPL * sizeof(typename PT::T),
                "a B fragment is EPL elements");
  const int n_frag = N * Kdim / PT::EPL;  // Kdim is a multiple of KSTEP
  typename PT::T* out = packed + pk_layer + off;
  const unsigned KB = (unsigned)(Kdim / PT::KSTEP);
  const float* Wi = thl + pick(tab.thw, i);  // W_i [h_i x in_i] (kinds 0 and 2)
  const int in_i = pick(tab.in_w, i), rows_i = pick(tab.rows, i);
  for (int q = blockIdx.x * 256 + threadIdx.x; q < n_frag; q += gridDim.x * 256) {
    const int lane = q & 63;
    const unsigned blk = (unsigned)q >> 6;
    const int nb = (int)(blk / KB), kb = (int)(blk - (unsigned)nb * KB);
    const int n = nb * 32 + (lane & 31);
    const int k0 = kb * PT::KSTEP + (lane >> 5) * PT::EPL;
    int64_t src = -1;  // kinds 0, 1, 4: offset of source row n
    if (kind == 0) {
      if (n < rows_i) src = (int64_t)n * in_i;
    } else if (kind == 1 || kind == 4) {
      const int row = kind == 1 ? dm.w3_row(l, n) : dm.w3r_row(l, n);
      if (row >= 0) src = (int64_t)row * hl;
    }
    typename PT::T el[PT::EPL];
#pragma unroll
    for (int e = 0; e < PT::EPL; ++e) {
      const int k = k0 + e;
      float v = 0.f;
      switch (kind) {
        case 0:  // row n of W_i, input column k (a conditioning dim for i = 0)
          if (src >= 0 && (i == 0 ? k < dm.dc(l) : k < in_i)) v = Wi[src + (i == 0 ? dm.cond(l, k) : k)];
          break;
        case 1:
        case 4:
          if (src >= 0 && k < hl) v = w3[src + k];
          break;
        case 2:  // transpose: row n = input column, k = output row of W_i
          if (k < rows_i && (i == 0 ? n < dm.dc(l) : n < in_i)) v = Wi[(int64_t)k * in_i + (i == 0 ? dm.cond(l, n) : n)];
          break;
        case 3: {
          const int row = dm.w3_row(l, k);
          if (row >= 0 && n < hl) v = w3[(int64_t)row * hl + n];
        } break;
        default: {
          const int row = dm.w3k_row(l, k);
          if (row >= 0 && n < hl) v = w3[(int64_t)row * hl + n];
        } break;
      }
      el[e] = PT::cvt(v);
    }
    typename PT::BF frag;
    __builtin_memcpy(&frag, el, sizeof(frag));
    *reinterpret_cast<typename PT::BF*>(out + (int64_t)q * PT::EPL) = frag;
  }
}

// ------------------------------------------------ base Gaussian factor ---
// chol(base_cov) for a full-covariance base Gaussian (Gaussian.log_prob /
// sample, common.py:285-293), one workgroup, column by column in global
// memory (left-looking, FP64 accumulation): L [d x d] row-major with zeros
// above the diagonal, then sum_j log L_jj. A covariance that is not
// positive definite gives NaN entries, as the reference's factorisation does.
__global__ void __launch_bounds__(256) base_chol_kernel(const float* __restrict__ cov, int d, float* L) {
  __shared__ double diag_s;
  for (int j = 0; j < d; ++j) {
    if (threadIdx.x == 0) {
      double s = cov[(int64_t)j * d + j];
      for (int k = 0; k < j; ++k) s -= (double)L[(int64_t)j * d + k] * L[(int64_t)j * d + k];
      diag_s = sqrt(s);
      L[(int64_t)j * d + j] = (float)diag_s;
    }
    __syncthreads();
    const double ljj = diag_s;
    for (int i = j + 1 + threadIdx.x; i < d; i += 256) {
      double s = cov[(int64_t)i * d + j];
      for (int k = 0; k < j; ++k) s -= (double)L[(int64_t)i * d + k] * L[(int64_t)j * d + k];
      L[(int64_t)i * d + j] = (float)(s / ljj);
    }
    for (int i = threadIdx.x; i < j; i += 256) L[(int64_t)i * d + j] = 0.f;  // upper triangle of column j
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double lg = 0.0;
    for (int j = 0; j < d; ++j) lg += log((double)L[(int64_t)j * d + j]);
    L[(int64_t)d * d] = (float)lg;
  }
}

// --------------------------------------------------------------- AdamW ---
__global__ void __launch_bounds__(256) sumsq_kernel(const float* __restrict__ g, int64_t n, double* part) {
  double s = 0.0;
  const int64_t n4 = n / 4;  // g is 16-byte aligned (checked by the caller)
  const float4* g4 = reinterpret_cast<const float4*>(g);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)gridDim.x * 256) {
    const float4 v = g4[i];
    s += (double)v.x * v.x;
    s += (double)v.y * v.y;
    s += (double)v.z * v.z;
    s += (double)v.w * v.w;
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
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

// b^e by square-and-multiply in FP64: the same sequence of products on the
// host (step given as an argument) and on the device (step read from device
// memory), so both give the same bias corrections bit for bit
__host__ __device__ inline double ipow_d(double b, int e) {
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

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
                                                    int64_t n, AdamHyper h, const double* nrm,
                                                    const int32_t* step_dev) {
  if (step_dev) {  // step count in device memory (a launch recorded once and replayed)
    const int t = *step_dev;
    h.bc1 = (float)(1.0 - ipow_d((double)h.b1, t));
    h.bc2 = (float)(1.0 - ipow_d((double)h.b2, t));
  }
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
  if (packed_bias_size) *packed_bias_size = dm.pb_alloc();
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
    for (int i = 0; i < dm.nh; ++i) {
      const int64_t c = (int64_t)dm.hp(i) * (dm.wh_k(l, i) > dm.wht_n(l, i) ? dm.wh_k(l, i) : dm.wht_n(l, i));
      mx = c > mx ? c : mx;
    }
    const int64_t c[3] = {(int64_t)dm.n3p(l) * dm.hlp(), (int64_t)dm.n3r(l) * dm.hlp(), dm.pb_size(l)};
    for (int64_t v : c) mx = v > mx ? v : mx;
  }
  FMC_ARG(mx < ((int64_t)1 << 31), "a packed weight image exceeds 2^31 elements");
  FMC_ARG((uintptr_t)fl->packed % 16 == 0, "packed must be 16-byte aligned");
  const PackTable tab = pack_table(dm);
  const int epl = fl->precision == FMC_PREC_BF16 ? 8 : 1;  // elements per B fragment
  for (int par = 0; par < 2; ++par) {
    FMC_ARG(tab.pk_size[par] % epl == 0, "packed layer size is not a whole number of fragments");
    for (int w = 0; w < tab.n_img; ++w)
      FMC_ARG(tab.off[par][w] % epl == 0, "packed image offset is not a whole number of fragments");
  }
  // one fragment per thread on the largest image up to 128 workgroups per
  // image: every workgroup decodes its image's table entries first, so a
  // larger grid costs more than it hides
  const int64_t nfx = (mx / epl + 255) / 256;
  unsigned gx = (unsigned)(nfx < 128 ? nfx : 128);
  dim3 grid(gx < 1 ? 1 : gx, (unsigned)((tab.n_img + 1) * dm.L));
  hipStream_t st = (hipStream_t)stream;
  if (fl->precision == FMC_PREC_BF16)
    hipLaunchKernelGGL(flow_pack_kernel<PBF16>, grid, dim3(256), 0, st, fd, tab, (__bf16*)fl->packed,
                       fl->packed_bias);
  else
    hipLaunchKernelGGL(flow_pack_kernel<PF32>, grid, dim3(256), 0, st, fd, tab, (float*)fl->packed, fl->packed_bias);
  int rc2 = check_launch("flow pack");
  if (rc2 || !dm.base_dense) return rc2;
  hipLaunchKernelGGL(base_chol_kernel, dim3(1), dim3(256), 0, st, fl->theta + dm.th_base_cov(), dm.d,
                     fl->packed_bias + dm.pb_base());
  return check_launch("base covariance factor");
}

extern "C" int fmc_flow_weight_grad(void* stream, const fmc_flow* fl, int32_t layer, const void* store,
                                    int64_t store_ld, const float* affine_partials, int32_t n_partials, float* grad) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(layer >= 0 && layer < fl->n_layers, "layer out of range");
  FMC_ARG(store && affine_partials && grad, "NULL pointer");
  FMC_ARG(store_ld >= 64 && store_ld % 64 == 0, "store_ld must be a positive multiple of 64");
  const WgTable tab = wgrad_table(fd.dm, layer);
  hipStream_t st = (hipStream_t)stream;
  const bool bf = fl->precision == FMC_PREC_BF16;
  FMC_ARG(n_partials >= store_ld / 32, "affine_partials holds %d slots, need %lld", n_partials,
          (long long)(store_ld / 32));
  const int per = (tab.total + WG_XCDS - 1) / WG_XCDS;
  dim3 g((unsigned)(per * WG_XCDS)), b(64 * WG_SPLIT);
  if (bf)
    hipLaunchKernelGGL(flow_wgrad_kernel<PBF16>, g, b, 0, st, fd, layer, tab, store, store_ld, affine_partials,
                       (int)(store_ld / 32), grad);
  else
    hipLaunchKernelGGL(flow_wgrad_kernel<PF32>, g, b, 0, st, fd, layer, tab, store, store_ld, affine_partials,
                       (int)(store_ld / 32), grad);
  return check_launch("flow weight grad");
}

// workgroups of the split weight gradient: one per two CUs. It runs beside
// the next layer's adjoint (313 workgroups at batch 10k, one or two per CU);
// a grid of one per CU delayed the adjoint (d=32, [128, 128]: 0.645 vs 0.527
// ms per training step; d=64, [256, 256]: 1.010 vs 1.022 ms)
static int wsplit_max_wg() { return device_cus() / 2; }

extern "C" int fmc_flow_weight_grad_ws_size(const fmc_flow* fl, int64_t store_ld, int64_t* bytes) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(bytes != nullptr, "bytes is NULL");
  FMC_ARG(store_ld >= 64 && store_ld % 64 == 0, "store_ld must be a positive multiple of 64");
  // the split sized for the whole chip (the last layer of a backward sweep,
  // when nothing runs beside it) needs the most
  int64_t mx = 0;
  for (int l = 0; l < (fd.dm.L < 2 ? fd.dm.L : 2); ++l) {  // layers alternate between two shapes
    for (int w : {wsplit_max_wg(), device_cus()}) {
      const WsTable t = wsplit_table(fd.dm, l, store_ld, w);
      const int64_t b = (int64_t)t.n_blocks * t.S * WS_PART * (int64_t)sizeof(float);
      mx = b > mx ? b : mx;
    }
  }
  *bytes = mx;
  return FMC_OK;
}

static int64_t wsplit_bytes(const FlowDims& dm, int layer, int64_t store_ld, int max_wg) {
  const WsTable t = wsplit_table(dm, layer, store_ld, max_wg);
  return (int64_t)t.n_blocks * t.S * WS_PART * (int64_t)sizeof(float);
}

// The split weight gradient with its batch ranges sized for max_wg
// workgroups in all (half the chip beside an adjoint, all of it alone).
static int weight_grad_split(void* stream, const fmc_flow* fl, int32_t layer, const void* store, int64_t store_ld,
                             const float* affine_partials, int32_t n_partials, float* grad, float* workspace,
                             int64_t workspace_bytes, int max_wg) {
  FlowDev fd;
  int rc = check_flow(fl, fd);
  if (rc) return rc;
  FMC_ARG(layer >= 0 && layer < fl->n_layers, "layer out of range");
  FMC_ARG(store && affine_partials && grad && workspace, "NULL pointer");
  FMC_ARG(store_ld >= 64 && store_ld % 64 == 0, "store_ld must be a positive multiple of 64");
  FMC_ARG(n_partials >= store_ld / 32, "affine_partials holds %d slots, need %lld", n_partials,
          (long long)(store_ld / 32));
  const WsTable tab = wsplit_table(fd.dm, layer, store_ld, max_wg);
  const int64_t need = (int64_t)tab.n_blocks * tab.S * WS_PART * (int64_t)sizeof(float);
  FMC_ARG(workspace_bytes >= need, "workspace holds %lld bytes, need %lld (fmc_flow_weight_grad_ws_size)",
          (long long)workspace_bytes, (long long)need);
  hipStream_t st = (hipStream_t)stream;
  const dim3 g((unsigned)(tab.n_blocks * tab.S)), b(64 * WS_WAVES);
  if (fl->precision == FMC_PREC_BF16) {
    const size_t shm = (size_t)2 * WS_KS * WS_FR * 64 * 16;
    set_lds(flow_wgrad_split_kernel<PBF16>, shm);
    hipLaunchKernelGGL(flow_wgrad_split_kernel<PBF16>, g, b, shm, st, fd, tab, store, store_ld, workspace);
  } else {
    const size_t shm = (size_t)2 * WS_KS * WS_FR * 64 * 4;
    set_lds(flow_wgrad_split_kernel<PF32>, shm);
    hipLaunchKernelGGL(flow_wgrad_split_kernel<PF32>, g, b, shm, st, fd, tab, store, store_ld, workspace);
  }
  rc = check_launch("flow weight grad (split)");
  if (rc) return rc;
  const int64_t ne = (int64_t)tab.n_blocks * WS_PART;
  // (one more workgroup than the elements need: the ScalarAffine sums)
  hipLaunchKernelGGL(flow_wgrad_reduce_kernel, dim3((unsigned)((ne + 255) / 256 + 1)), dim3(256), 0, st, fd, layer,
                     tab, workspace, affine_partials, (int)(store_ld / 32), grad);
  return check_launch("flow weight grad (reduce)");
}

extern "C" int fmc_flow_weight_grad_split(void* stream, const fmc_flow* fl, int32_t layer, const void* store,
                                          int64_t store_ld, const float* affine_partials, int32_t n_partials,
                                          float* grad, float* workspace, int64_t workspace_bytes) {
  return weight_grad_split(stream, fl, layer, store, store_ld, affine_partials, n_partials, grad, workspace,
                           workspace_bytes, wsplit_max_wg());
}

static int adamw_clip(void* stream, float* theta, const float* grad, float* m, float* v, int64_t n, float lr,
                      float b1, float b2, float eps, float weight_decay, float max_norm, int32_t step,
                      const int32_t* step_dev, double* workspace, int32_t workspace_len, float* grad_norm) {
  FMC_ARG(theta && grad && m && v && workspace, "NULL pointer");
  FMC_ARG(n >= 0 && (step_dev != nullptr || step >= 1), "n must be >= 0 and step >= 1");
  FMC_ARG(workspace_len >= 2, "workspace_len must be >= 2");
  FMC_ARG(((uintptr_t)theta | (uintptr_t)grad | (uintptr_t)m | (uintptr_t)v) % 16 == 0,
          "theta, grad, m and v must be 16-byte aligned");
  if (n == 0) return FMC_OK;
  int nb = (int)((n / 4 + 255) / 256);
  if (nb < 1) nb = 1;
  if (nb > workspace_len - 1) nb = workspace_len - 1;
  if (nb > 1024) nb = 1024;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(sumsq_kernel, dim3(nb), dim3(256), 0, st, grad, n, workspace);
  int rc = check_launch("adamw sumsq");
  if (rc) return rc;
  hipLaunchKernelGGL(norm_kernel, dim3(1), dim3(256), 0, st, workspace, nb, grad_norm);
  if ((rc = check_launch("adamw norm"))) return rc;
  const int t = step_dev ? 1 : step;  // (the kernel recomputes the corrections from step_dev)
  AdamHyper h{lr, b1, b2, eps, weight_decay, max_norm, (float)(1.0 - ipow_d((double)b1, t)),
              (float)(1.0 - ipow_d((double)b2, t))};
  int64_t nb2 = (n / 4 + 255) / 256;  // one float4 per thread up to 2M parameters
  if (nb2 > 8192) nb2 = 8192;
  if (nb2 < 1) nb2 = 1;
  hipLaunchKernelGGL(adamw_kernel, dim3((unsigned)nb2), dim3(256), 0, st, theta, grad, m, v, n, h, workspace + nb, step_dev);
  return check_launch("adamw update");
}

extern "C" int fmc_adamw_clip(void* stream, float* theta, const float* grad, float* m, float* v, int64_t n, float lr,
                              float b1, float b2, float eps, float weight_decay, float max_norm, int32_t step,
                              double* workspace, int32_t workspace_len, float* grad_norm) {
  return adamw_clip(stream, theta, grad, m, v, n, lr, b1, b2, eps, weight_decay, max_norm, step, nullptr, workspace,
                    workspace_len, grad_norm);
}

extern "C" int fmc_adamw_clip_dev(void* stream, float* theta, const float* grad, float* m, float* v, int64_t n,
                                  float lr, float b1, float b2, float eps, float weight_decay, float max_norm,
                                  const int32_t* step_dev, double* workspace, int32_t workspace_len,
                                  float* grad_norm) {
  FMC_ARG(step_dev != nullptr, "step_dev is NULL");
  return adamw_clip(stream, theta, grad, m, v, n, lr, b1, b2, eps, weight_decay, max_norm, 0, step_dev, workspace,
                    workspace_len, grad_norm);
}

// The whole backward sweep of a training step (nf_model/base.py:98-125, the
// gradient of NFModel.loss_fn): per layer, last to first, the adjoint on
// `stream`, then that layer's weight gradient. With a side stream the weight
// gradient of layer l runs there beside the adjoint of layer l-1: the two
// stores / affine-partial buffers alternate between layers, an adjoint waits
// for the weight gradient that last read its store, and `stream` joins the
// side stream before returning. The events are the call's own (created and
// released here), so overlapping calls on other streams do not interfere.
// p[0:n) = 0 (a kernel rather than a memset, so that a caller may record the
// sweep into a HIP graph); float4 stores when p is 16-byte aligned
__global__ void __launch_bounds__(256) zero_kernel(float* __restrict__ p, int64_t n, int vec) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x, stride = (int64_t)gridDim.x * 256;
  int64_t i0 = t;
  if (vec) {
    float4* p4 = reinterpret_cast<float4*>(p);
    const int64_t n4 = n / 4;
    for (int64_t i = t; i < n4; i += stride) p4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    i0 = 4 * n4 + t;
  }
  for (int64_t i = i0; i < n; i += stride) p[i] = 0.f;
}

extern "C" int fmc_flow_backward(void* stream, void* side_stream, const fmc_flow* fl, const float* saved, int64_t n,
                                 const float* g_logp, float* gy, float* gx, void* store0, void* store1,
                                 int64_t store_ld, float* aff0, float* aff1, int32_t n_partials, float* grad,
                                 int32_t split, float* wg_workspace, int64_t wg_workspace_bytes) {
  FMC_ARG(fl != nullptr && saved && g_logp && gy && gx && store0 && aff0 && grad, "NULL pointer");
  const bool two = side_stream != nullptr;
  FMC_ARG(!two || (store1 != nullptr && aff1 != nullptr), "the side-stream schedule needs store1 and aff1");
  FMC_ARG(!split || wg_workspace != nullptr, "the split weight gradient needs its workspace");
  FMC_ARG(fl->n_layers >= 1 && n >= 1, "bad n_layers / n");
  FlowDev fd;
  if (const int rcf = check_flow(fl, fd)) return rcf;
  hipStream_t main_st = (hipStream_t)stream, side = two ? (hipStream_t)side_stream : main_st;
  auto hip_fail = [](const char* what) {
    set_error(FMC_ERR_HIP, "%s failed", what);
    return (int)FMC_ERR_HIP;
  };
  hipEvent_t ev_adj = nullptr, ev_wg[2] = {nullptr, nullptr};
  if (two) {
    for (hipEvent_t* e : {&ev_adj, &ev_wg[0], &ev_wg[1]}) {
      if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
        for (hipEvent_t x : {ev_adj, ev_wg[0], ev_wg[1]})
          if (x) (void)hipEventDestroy(x);
        return hip_fail("hipEventCreateWithFlags");
      }
    }
  }
  bool pending[2] = {false, false};
  int rc = FMC_OK;
  const int L = fl->n_layers;
  const int64_t layer_elems = n * (int64_t)fl->n_dim;
  float *gin = gy, *gout = gx;
  for (int i = 0; i < L; ++i) {
    const int layer = L - 1 - i, b = two ? (i & 1) : 0;
    void* st = b ? store1 : store0;
    float* af = b ? aff1 : aff0;
    if (pending[b] && hipStreamWaitEvent(main_st, ev_wg[b], 0) != hipSuccess) {
      rc = hip_fail("hipStreamWaitEvent");
      break;
    }
    rc = fmc_flow_backward_layer(stream, fl, layer, saved + layer * layer_elems, n, g_logp, gin, gout, st, store_ld, af,
                                 n_partials);
    if (rc) break;
    if (two && (hipEventRecord(ev_adj, main_st) != hipSuccess || hipStreamWaitEvent(side, ev_adj, 0) != hipSuccess)) {
      rc = hip_fail("hipEventRecord / hipStreamWaitEvent");
      break;
    }
    if (i == 0) {  // clear grad where the weight gradients are written, ahead of the first
      const int64_t nz = fd.dm.th_total();
      const int vec = ((uintptr_t)grad % 16) == 0;
      const int64_t nzb = ((vec ? nz / 4 : nz) + 255) / 256;
      hipLaunchKernelGGL(zero_kernel, dim3((unsigned)(nzb < 1 ? 1 : nzb > 2048 ? 2048 : nzb)), dim3(256), 0, side, grad,
                         nz, vec);
      if ((rc = check_launch("gradient clear"))) break;
    }
    if (split) {
      // the last layer's weight gradient runs alone: batch ranges for the
      // whole chip when the workspace holds them (the same choice on one
      // stream or two, so the schedule does not change the sums)
      int max_wg = wsplit_max_wg();
      if (layer == 0 && wsplit_bytes(fd.dm, 0, store_ld, device_cus()) <= wg_workspace_bytes) max_wg = device_cus();
      rc = weight_grad_split(side, fl, layer, st, store_ld, af, n_partials, grad, wg_workspace, wg_workspace_bytes,
                             max_wg);
    } else {
      rc = fmc_flow_weight_grad(side, fl, layer, st, store_ld, af, n_partials, grad);
    }
    if (rc) break;
    if (two) {
      if (hipEventRecord(ev_wg[b], side) != hipSuccess) {
        rc = hip_fail("hipEventRecord");
        break;
      }
      pending[b] = true;
    }
    float* t = gin;
    gin = gout;
    gout = t;
  }
  if (two) {
    for (int b = 0; b < 2; ++b)
      if (pending[b] && hipStreamWaitEvent(main_st, ev_wg[b], 0) != hipSuccess && rc == FMC_OK)
        rc = hip_fail("hipStreamWaitEvent");
    for (hipEvent_t x : {ev_adj, ev_wg[0], ev_wg[1]}) (void)hipEventDestroy(x);
  }
  return rc;
}
