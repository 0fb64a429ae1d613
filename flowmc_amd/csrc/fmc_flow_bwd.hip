// Adjoint of one MaskedCouplingRQSpline coupling layer on gfx950 (the
// autodiff of NFModel.loss_fn, nf_model/base.py:98-100, done by hand):
// recomputes the conditioner and spline parameters, runs the spline VJP in
// registers and the transposed conditioner GEMMs, and stores activations and
// gradients for the weight-gradient kernel (fragment-blocked).
#include "fmc_flow_common.h"

namespace fmc {

// Tile of the fragment-blocked store: rows row0 .. row0+nf (row0 and nf
// multiples of 32) x samples m0 .. m0+MT, from src[m * lds + f] (rows f >=
// nvalid are written as zeros). One thread per EPL-element lane chunk, so
// consecutive threads write consecutive chunks of a 64-lane block.
template <class PT, int MT>
__device__ __forceinline__ void store_fm(typename PT::T* st, int64_t ld, int row0, int nf, int nvalid,
                                         const typename PT::T* src, int lds, int64_t m0) {
  using T = typename PT::T;
  constexpr int EPL = PT::EPL, KS = PT::KSTEP;
  constexpr int SPT = MT / KS;  // batch slices in the tile
  const int64_t ns = ld / KS;
  const int nchunk = (nf / 32) * SPT * 64;
  for (int c = threadIdx.x; c < nchunk; c += 256) {
    const int lane = c & 63, sl = (c >> 6) % SPT, rbl = (c >> 6) / SPT;
    const int f = rbl * 32 + (lane & 31);
    const int mb = sl * KS + (lane >> 5) * EPL;
    T* dst = st + (((int64_t)(row0 / 32 + rbl) * ns + m0 / KS + sl) * 64 + lane) * EPL;
    if constexpr (EPL * sizeof(T) == 16) {  // one 16-byte store per lane chunk
      union {
        T t[EPL];
        uint4 u;
      } pk;
      if (f < nvalid) {
#pragma unroll
        for (int e = 0; e < EPL; ++e) pk.t[e] = src[(mb + e) * lds + f];
      } else {
        pk.u = uint4{0u, 0u, 0u, 0u};
      }
      *reinterpret_cast<uint4*>(dst) = pk.u;
    } else {
#pragma unroll
      for (int e = 0; e < EPL; ++e) dst[e] = (f < nvalid) ? src[(mb + e) * lds + f] : PT::cvt(0.f);
    }
  }
}

// 1 - h^2 of one lane's 16 elements of a 32x32 hidden block, kept in
// registers from the forward recompute to the adjoint (computed in FP32 from
// the FP32 activation: the BF16 copy of h in LDS would wreck it near
// saturation). BF16 conditioner: stored as BF16 pairs (relative rounding
// 2^-9, below the BF16 rounding of the gradients it scales), which frees the
// registers of a deeper weight-prefetch ring; FP32 conditioner: FP32.
template <class PT>
struct DhBlock {
  float v[16];
  __device__ __forceinline__ void put2(int i, float a, float b) {
    v[2 * i] = a;
    v[2 * i + 1] = b;
  }
  __device__ __forceinline__ float get(int i) const { return v[i]; }
};
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
template <>
struct DhBlock<PBF16> {
  bf16x2_t p[8];
  __device__ __forceinline__ void put2(int i, float a, float b) { p[i] = bf16x2_t{(__bf16)a, (__bf16)b}; }
  __device__ __forceinline__ float get(int i) const { return (float)p[i >> 1][i & 1]; }
};

template <class PT, int MB, int HMAX, int KT, int NH>
__device__ __forceinline__ void flow_bwd_body(const FlowDev& f, const BwdArgs& a) {
  using T = typename PT::T;
  constexpr int MT = 32 * MB;
  constexpr int P = BWD_PF;     // ring depth of the [MT x hidden] products
  constexpr int PR = BWD_PF_R;  // ring depth of the transposed output product (two fragments a step)
  // 32x32 tiles per wave of an [MT x hidden] block (hidden <= HMAX); the same
  // wave owns the same tiles in the forward recompute and in the adjoint, so
  // the tanh derivatives of all NH hidden layers stay in its registers.
  constexpr int NCW = (HMAX / 32) / 4;  // column blocks per wave (even)
  extern __shared__ __align__(16) unsigned char smem[];
  const FlowDims& dm = f.dm;
  BwdSmem<PT> sm(dm, MT);
  float* xs = reinterpret_cast<float*>(smem + sm.xs);
  float* gb = reinterpret_cast<float*>(smem + sm.gb);
  float* glp = reinterpret_cast<float*>(smem + sm.glp);
  // hidden layer i (and later its pre-activation gradient) lives in hbuf[i & 1]
  T* const hbuf[2] = {reinterpret_cast<T*>(smem + sm.hb1), reinterpret_cast<T*>(smem + sm.hb2)};
  T* const hlast = hbuf[(NH - 1) & 1];
  float* obuf = reinterpret_cast<float*>(smem + sm.ob);
  T* gpb = reinterpret_cast<T*>(smem + sm.gpb);
  float* red = reinterpret_cast<float*>(smem + sm.red);
  // the wave index in a scalar register: branches on it (column blocks,
  // dim pairs) are uniform, not exec-masked
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int d = dm.d, ldx = sm.ldx, l = a.layer;
  const int64_t m0 = (int64_t)blockIdx.x * MT;
  const float* th = f.theta;
  const T* W = reinterpret_cast<const T*>(f.packed) + dm.pk_layer(l);
  const float* bl = f.pbias + dm.pb_layer(l);
  T* st = reinterpret_cast<T*>(a.store);
  const int hlp = dm.hlp(), P3 = dm.P();
  const float sc = th[dm.th_scale(l)], sh = th[dm.th_shift(l)];
  const float es = expf(sc);

  // ---- load x (layer input), gy, g_logp: a thread issues all its loads of
  // a batch of 8 elements before it writes any of them, so the tile costs
  // one memory round trip per 2048 elements, not one per 256
  {
    const int tot = MT * d;
    const int64_t base = m0 * d, lim = (a.n - m0) * (int64_t)d;  // valid elements of the tile
    const float gl = (tid < MT && m0 + tid < a.n) ? a.g_logp[m0 + tid] : 0.f;
    for (int e0 = 0; e0 < tot; e0 += 256 * 8) {
      float xv[8], gv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * 256 + tid;
        const bool ok = e < tot && e < lim;
        xv[u] = ok ? a.x_in[base + e] : 0.f;
        gv[u] = ok ? a.gy[base + e] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * 256 + tid;
        if (e < tot) {
          const int m = e / d, j = e - m * d;
          xs[m * ldx + j] = (e < lim) ? (xv[u] + sh) * es : 0.f;
          gb[m * ldx + j] = gv[u];
        }
      }
    }
    if (tid < MT) glp[tid] = gl;
  }
  __syncthreads();
  // ---- the conditioner input (kc columns, zero beyond dc)
  const int dc = dm.dc(l), kc = dm.kc(l);
  T* a0 = gpb;
  const int la0 = sm.lg;
  for (TileWalk tw(tid, kc); tw.m < MT; tw.next()) {
    const int m = tw.m, k = tw.j;
    a0[m * la0 + k] = PT::cvt(k < dc ? xs[m * ldx + dm.cond(l, k)] : 0.f);
  }
  // Stores of the fragment-blocked activations / gradients are issued after
  // the MFMAs of the next product (the LDS copy is still intact then), so
  // their completion overlaps the epilogue instead of delaying the next
  // weight fragments (loads and stores share one in-order counter).
  // ---- recompute the conditioner forward; wave w owns column blocks w,
  // w+4, w+8, ... (all MB row blocks) of every [MT x hidden] product; index
  // ci of the register arrays is the wave-local column block.
  DhBlock<PT> dh[NH][NCW][MB];
  {
    const T* in = a0;
    int ldin = la0, KB = kc / PT::KSTEP;
#pragma unroll
    for (int i = 0; i < NH; ++i) {
      T* out = hbuf[i & 1];
      const int NBt = dm.hp(i) / 32;
      const float* bi = bl + dm.pb_b(i);
      const T* Wi = W + dm.pk_wh(l, i);
      BRing<PT, 2, P> ring;
      if (wave < NBt) panel_prefetch(ring, Wi, KB, wave, 4, (wave + 4 < NBt) ? 2 : 1, 0, KB, lane);
      __syncthreads();  // `in` complete
#pragma unroll
      for (int ci = 0; ci < NCW; ci += 2) {
        const int c0 = wave + 4 * ci;
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        if (c0 < NBt) {
          if (ci > 0) panel_prefetch(ring, Wi, KB, c0, 4, ncb, 0, KB, lane);
          panel_run<PT, MB, 2, P>(ring, in, ldin, Wi, KB, c0, 4, ncb, 0, KB, 0, acc, lane);
        }
        if (ci == 0) {  // the previous product's input
          if (i == 0)
            store_fm<PT, MT>(st, a.ld, (int)dm.st_c(l), dm.ncp(l), kc, a0, la0, m0);
          else
            store_fm<PT, MT>(st, a.ld, (int)dm.st_h(l, i - 1), dm.hp(i - 1), dm.hp(i - 1), in, sm.lh, m0);
        }
        if (c0 < NBt) {
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if (j < ncb) {
              const int col = (c0 + 4 * j) * 32 + (lane & 31);
              const float bb = bi[col];
#pragma unroll
              for (int r = 0; r < MB; ++r)
#pragma unroll
                for (int v = 0; v < 16; v += 2) {
                  float h0, h1;
                  PT::act2(acc[j][r][v] + bb, acc[j][r][v + 1] + bb, h0, h1);
                  out[(r * 32 + out_row(v, lane)) * sm.lh + col] = PT::cvt(h0);
                  out[(r * 32 + out_row(v + 1, lane)) * sm.lh + col] = PT::cvt(h1);
                  dh[i][ci + j][r].put2(v >> 1, 1.f - h0 * h0, 1.f - h1 * h1);
                }
            }
          }
        }
      }
      in = out;
      ldin = sm.lh;
      KB = dm.hp(i) / PT::KSTEP;
    }
  }

  // ---- chunks: params -> spline VJP -> gp; g_hl += gp W3c (register
  // accumulators over the columns of the last hidden layer)
  f32x16_t gh2[NCW][MB];
#pragma unroll
  for (int ci = 0; ci < NCW; ++ci)
#pragma unroll
    for (int r = 0; r < MB; ++r) gh2[ci][r] = f32x16_t{};
  const int NB2 = hlp / 32;
  if (dm.reg_spline()) {
    // rounds of 4 dim pairs (one per wave): transposed GEMM3 -> parameters in
    // registers -> spline VJP in registers -> gp (gradient order) in LDS ->
    // g_hl += gp W3. Rounds alternate between two gp buffers (gpb and the
    // hidden buffer not holding the last hidden layer, free by now), so a
    // round needs one barrier; a round's gp is stored at its end, after the
    // next round's first W3 fragments are requested.
    // (selected as LDS byte offsets: a pointer picked from an array lost its
    // address space and became FLAT accesses, which the loads' waits count)
    const int gp_off[2] = {sm.gpb, (NH & 1) ? sm.hb2 : sm.hb1};
    const int npair = dm.n_pairs(l), dtl = dm.dt(l), K = dm.K;
    const T* W3r = W + dm.pk_w3r(l);
    const T* W3rt = W + dm.pk_w3rt(l);
    const float* b3r = bl + dm.pb_b3r(l);
    const int KB3 = hlp / PT::KSTEP, KB3t = dm.n3r(l) / PT::KSTEP;
    const int hf = lane >> 5;
    BRing<PT, 2, PR> ring;
    if (wave < npair) panel_prefetch(ring, W3r, KB3, 2 * wave, 1, 2, 0, KB3, lane);
    __syncthreads();  // the last hidden layer complete
    store_fm<PT, MT>(st, a.ld, (int)dm.st_h(l, NH - 1), hlp, hlp, hlast, sm.lh, m0);
    int rnd = 0;
    for (int q0 = 0; q0 < npair; q0 += 4, ++rnd) {
      const int nq = (npair - q0) < 4 ? (npair - q0) : 4;
      T* const gcur = reinterpret_cast<T*>(smem + ((rnd & 1) ? gp_off[1] : gp_off[0]));
      const int q = q0 + wave;
      if (wave < nq) {
        f32x16_t acc[2][MB] = {};
        panel_run<PT, MB, 2, PR, true>(ring, hlast, sm.lh, W3r, KB3, 2 * q, 1, 2, 0, KB3, 0, acc, lane);
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
            const float gxv = rs_vjp<KT, sizeof(T) == 2>(pp, K, f, xs[m * ldx + j], gb[m * ldx + j], glp[m], go);
            gb[m * ldx + j] = gxv;
          } else {
#pragma unroll
            for (int i = 0; i < 32; ++i) go[i] = 0.f;
          }
          // slot 8t+e of this lane's dim -> gradient-order column 64 wave + 16t + 8hf + e
          T* dst = gcur + m * sm.lg + 64 * wave + 8 * hf;
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
      // g_hl += gp W3 over this round's k-range
      const int kb0 = 64 * q0 / PT::KSTEP, kb1 = 64 * (q0 + nq) / PT::KSTEP;
      const int c0 = wave;  // column blocks wave, wave + 4 (NCW == 2) or wave + 4*ci
      BRing<PT, 2, P> ring2;
      if (c0 < NB2) panel_prefetch(ring2, W3rt, KB3t, c0, 4, (c0 + 4 < NB2) ? 2 : 1, kb0, kb1, lane);
      __syncthreads();  // gcur complete
#pragma unroll
      for (int ci = 0; ci < NCW; ci += 2) {
        const int cb = wave + 4 * ci;
        if (cb < NB2) {
          const int ncb = (cb + 4 < NB2) ? 2 : 1;
          if (ci > 0) panel_prefetch(ring2, W3rt, KB3t, cb, 4, ncb, kb0, kb1, lane);
          f32x16_t acc2[2][MB];
#pragma unroll
          for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < MB; ++r) acc2[jj][r] = gh2[ci + jj][r];
          panel_run<PT, MB, 2, P>(ring2, gcur, sm.lg, W3rt, KB3t, cb, 4, ncb, kb0, kb1, 0, acc2, lane);
#pragma unroll
          for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < MB; ++r) gh2[ci + jj][r] = acc2[jj][r];
        }
      }
      // this wave's pair in the next round (the ring is always rewritten, so
      // its old contents are dead during the VJP)
      const int qn = q0 + 4 + wave;
      if (qn < npair) {
        panel_prefetch(ring, W3r, KB3, 2 * qn, 1, 2, 0, KB3, lane);
      } else {
        ring = BRing<PT, 2, PR>{};
      }
      store_fm<PT, MT>(st, a.ld, (int)dm.st_gp(l) + 64 * q0, 64 * nq, 64 * nq, gcur, sm.lg, m0);
    }
    // every read of the last hidden layer ended before the last round's
    // barrier; the gp buffers are next written after a barrier
  } else {
    __syncthreads();  // the last hidden layer complete
    store_fm<PT, MT>(st, a.ld, (int)dm.st_h(l, NH - 1), hlp, hlp, hlast, sm.lh, m0);
    const int nch = dm.n_chunks(l);
    const int KB3 = hlp / PT::KSTEP;
    const int KB3t = dm.n3p(l) / PT::KSTEP;
    for (int c = 0; c < nch; ++c) {
      const int nc = dm.chunk_nc(l, c), cnt = dm.chunk_cnt(l, c), coff = dm.chunk_off(l, c);
      const int NBt = nc / 32;
      for (int c0 = wave; c0 < NBt; c0 += 8) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, PF_DEPTH>(hlast, sm.lh, W + dm.pk_w3(l), KB3, coff / 32 + c0, 4, ncb, 0, KB3, 0, acc,
                                        lane);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < ncb) {
            const int col = (c0 + 4 * j) * 32 + (lane & 31);
            const float bb = bl[dm.pb_b3(l) + coff + col];
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
        const float* pp = obuf + m * sm.lo_ + jj * P3;
        float gyv = gb[m * ldx + j];
        float gxv = spline_vjp<T>(pp, f, xs[m * ldx + j], gyv, glp[m], gpb + m * sm.lg + jj * P3,
                                  [](float v) { return PT::cvt(v); });
        gb[m * ldx + j] = gxv;
      }
      for (int i = tid; i < MT * (nc - cnt * P3); i += 256) {
        int w = nc - cnt * P3;
        int m = i / w, cc = cnt * P3 + (i - m * w);
        gpb[m * sm.lg + cc] = PT::cvt(0.f);
      }
      __syncthreads();
      store_fm<PT, MT>(st, a.ld, (int)dm.st_gp(l) + coff, nc, nc, gpb, sm.lg, m0);
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
  }
  // ---- g_{NH-1} = g_hl * (1 - h_{NH-1}^2)   (in place of h_{NH-1}: every
  // read of it ended before the last barrier)
#pragma unroll
  for (int ci = 0; ci < NCW; ++ci) {
    const int nb = wave + 4 * ci;
    if (nb < NB2) {
      const int col = nb * 32 + (lane & 31);
#pragma unroll
      for (int r = 0; r < MB; ++r)
#pragma unroll
        for (int v = 0; v < 16; ++v)
          hlast[(r * 32 + out_row(v, lane)) * sm.lh + col] = PT::cvt(gh2[ci][r][v] * dh[NH - 1][ci][r].get(v));
    }
  }
  // ---- g_{i-1} = (g_i W_i) * (1 - h_{i-1}^2), i = NH-1 .. 1   (in place of h_{i-1})
#pragma unroll
  for (int i = NH - 1; i >= 1; --i) {
    const T* gin = hbuf[i & 1];
    T* gout = hbuf[(i - 1) & 1];
    const int KB = dm.hp(i) / PT::KSTEP, NBt = dm.hp(i - 1) / 32;
    const T* Wt = W + dm.pk_wht(l, i);
    BRing<PT, 2, P> ring;
    if (wave < NBt) panel_prefetch(ring, Wt, KB, wave, 4, (wave + 4 < NBt) ? 2 : 1, 0, KB, lane);
    __syncthreads();  // g_i complete
#pragma unroll
    for (int ci = 0; ci < NCW; ci += 2) {
      const int c0 = wave + 4 * ci;
      const int ncb = (c0 + 4 < NBt) ? 2 : 1;
      f32x16_t acc[2][MB] = {};
      if (c0 < NBt) {
        if (ci > 0) panel_prefetch(ring, Wt, KB, c0, 4, ncb, 0, KB, lane);
        panel_run<PT, MB, 2, P>(ring, gin, sm.lh, Wt, KB, c0, 4, ncb, 0, KB, 0, acc, lane);
      }
      if (ci == 0) store_fm<PT, MT>(st, a.ld, (int)dm.st_gh(l, i), dm.hp(i), dm.hp(i), gin, sm.lh, m0);
      if (c0 < NBt) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < ncb) {
            const int col = (c0 + 4 * j) * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v)
                gout[(r * 32 + out_row(v, lane)) * sm.lh + col] =
                    PT::cvt(acc[j][r][v] * dh[i - 1][ci + j][r].get(v));
          }
        }
      }
    }
  }
  // ---- g_c = g_0 W_0c, scattered onto the conditioning dims. With one
  // column block (dc <= 32) the k-range is split over the four waves and
  // their partial blocks summed in a fixed order (in the gp buffer, free now).
  {
    const int KB = dm.hp(0) / PT::KSTEP, NBt = dm.ncp(l) / 32;
    const T* W0t = W + dm.pk_wht(l, 0);
    const bool split = NBt == 1 && 4 * MT * 33 * 4 <= MT * sm.lg * (int)sizeof(T);
    if (split) {
      const int k0 = wave * KB / 4, k1 = (wave + 1) * KB / 4;
      BRing<PT, 2, P> ring;
      panel_prefetch(ring, W0t, KB, 0, 4, 1, k0, k1, lane);
      __syncthreads();  // g_0 complete
      f32x16_t acc[2][MB] = {};
      panel_run<PT, MB, 2, P>(ring, hbuf[0], sm.lh, W0t, KB, 0, 4, 1, k0, k1, k0, acc, lane);
      store_fm<PT, MT>(st, a.ld, (int)dm.st_gh(l, 0), dm.hp(0), dm.hp(0), hbuf[0], sm.lh, m0);
      float* part = reinterpret_cast<float*>(gpb);
#pragma unroll
      for (int r = 0; r < MB; ++r)
#pragma unroll
        for (int v = 0; v < 16; ++v) part[(wave * MT + r * 32 + out_row(v, lane)) * 33 + (lane & 31)] = acc[0][r][v];
      __syncthreads();
      for (int e = tid; e < MT * 32; e += 256) {
        const int m = e >> 5, k = e & 31;
        if (k < dc) {
          const float s4 = ((part[m * 33 + k] + part[(MT + m) * 33 + k]) + part[(2 * MT + m) * 33 + k]) +
                           part[(3 * MT + m) * 33 + k];
          gb[m * ldx + dm.cond(l, k)] += s4;
        }
      }
    } else {
      __syncthreads();  // g_0 complete
      store_fm<PT, MT>(st, a.ld, (int)dm.st_gh(l, 0), dm.hp(0), dm.hp(0), hbuf[0], sm.lh, m0);
      for (int c0 = wave; c0 < NBt; c0 += 8) {
        const int ncb = (c0 + 4 < NBt) ? 2 : 1;
        f32x16_t acc[2][MB] = {};
        mfma_panel<PT, MB, 2, P>(hbuf[0], sm.lh, W0t, KB, c0, 4, ncb, 0, KB, 0, acc, lane);
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

// 32-sample tiles. Hidden widths up to 256 with one or two hidden layers:
// register budget held to two waves per SIMD (no spills), and with ~68 KB
// of LDS per tile two workgroups share a CU. Three or four hidden layers
// keep their tanh derivatives in registers too and take the default budget
// (one wave per SIMD, VGPRs and AGPRs), as do hidden widths up to 512
// (one or two hidden layers).
template <class PT, int KT, int NH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
flow_bwd_kernel_h256(FlowDev f, BwdArgs a) {
  flow_bwd_body<PT, 1, 256, KT, NH>(f, a);
}

template <class PT, int KT, int NH>
__global__ void __launch_bounds__(256) flow_bwd_kernel_h256_deep(FlowDev f, BwdArgs a) {
  flow_bwd_body<PT, 1, 256, KT, NH>(f, a);
}

template <class PT, int KT, int NH>
__global__ void __launch_bounds__(256) flow_bwd_kernel_h512(FlowDev f, BwdArgs a) {
  flow_bwd_body<PT, 1, 512, KT, NH>(f, a);
}

template <class PT, int KT, int NH>
static void launch_bwd(bool wide, dim3 g, dim3 b, size_t shm, hipStream_t st, const FlowDev& fd, const BwdArgs& a) {
  if constexpr (NH <= 2) {
    if (wide) {
      set_lds(flow_bwd_kernel_h512<PT, KT, NH>, shm);
      hipLaunchKernelGGL((flow_bwd_kernel_h512<PT, KT, NH>), g, b, shm, st, fd, a);
    } else {
      set_lds(flow_bwd_kernel_h256<PT, KT, NH>, shm);
      hipLaunchKernelGGL((flow_bwd_kernel_h256<PT, KT, NH>), g, b, shm, st, fd, a);
    }
  } else {
    set_lds(flow_bwd_kernel_h256_deep<PT, KT, NH>, shm);
    hipLaunchKernelGGL((flow_bwd_kernel_h256_deep<PT, KT, NH>), g, b, shm, st, fd, a);
  }
}

template <class PT, int KT>
static void launch_bwd_nh(int nh, bool wide, dim3 g, dim3 b, size_t shm, hipStream_t st, const FlowDev& fd,
                          const BwdArgs& a) {
  switch (nh) {
    case 1: launch_bwd<PT, KT, 1>(wide, g, b, shm, st, fd, a); break;
    case 2: launch_bwd<PT, KT, 2>(wide, g, b, shm, st, fd, a); break;
    case 3: launch_bwd<PT, KT, 3>(wide, g, b, shm, st, fd, a); break;
    default: launch_bwd<PT, KT, 4>(wide, g, b, shm, st, fd, a); break;
  }
}

template <class PT>
static void launch_bwd_kt(int kt, int nh, bool wide, dim3 g, dim3 b, size_t shm, hipStream_t st, const FlowDev& fd,
                          const BwdArgs& a) {
  if (kt == 8)
    launch_bwd_nh<PT, 8>(nh, wide, g, b, shm, st, fd, a);
  else if (kt == 10)
    launch_bwd_nh<PT, 10>(nh, wide, g, b, shm, st, fd, a);
  else
    launch_bwd_nh<PT, 0>(nh, wide, g, b, shm, st, fd, a);
}

}  // namespace fmc

using namespace fmc;

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
  const int hmax = fd.dm.hmaxp();
  FMC_REQUIRE(hmax <= 512, FMC_ERR_UNSUPPORTED, "hidden sizes > 512 not supported by the backward kernel");
  FMC_REQUIRE(hmax <= 256 || fd.dm.nh <= 2, FMC_ERR_UNSUPPORTED,
              "the backward kernel takes hidden sizes above 256 with at most two hidden layers");
  // every store column [0, store_ld) is rewritten (rows >= n carry zero gradients)
  const int64_t nblk = store_ld / 32;
  FMC_ARG(n_partials >= nblk, "affine_partials holds %d slots, need %lld", n_partials, (long long)nblk);
  BwdArgs a{layer, x_layer, n, g_logp, gy, gx, store, store_ld, affine_partials};
  dim3 g((unsigned)nblk), b(256);
  hipStream_t st = (hipStream_t)stream;
  const int kt = spline_kt(fd.dm.K);
  if (bf)
    launch_bwd_kt<PBF16>(kt, fd.dm.nh, hmax > 256, g, b, shm, st, fd, a);
  else
    launch_bwd_kt<PF32>(kt, fd.dm.nh, hmax > 256, g, b, shm, st, fd, a);
  return check_launch("flow backward");
}

