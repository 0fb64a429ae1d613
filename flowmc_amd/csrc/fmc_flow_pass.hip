// MaskedCouplingRQSpline flow passes on gfx950: all coupling layers of a
// forward / inverse pass for a tile of samples in one workgroup, the
// conditioner GEMMs on MFMA with activations in LDS and the spline fused as
// the epilogue of the last GEMM (rqSpline.py:450-504).
#include "fmc_flow_common.h"

namespace fmc {

// 32-sample tiles (MB = 1) with the register budget held to three waves per
// SIMD: three workgroups share a CU (LDS ~43 KB each at d=64, [256,256]) and
// hide each other's LDS / L2 latency; measured faster than 64-sample tiles
// with more B-fragment reuse per wave.
// One panel of the pass: EX (grids of at most two workgroups per CU, e.g.
// training batches of 10^4 samples) uses the exact-wait ring, four k-steps
// deep, which pays where a wave has few neighbours to hide its waits; the
// guarded 8-deep ring otherwise (sampling passes of 10^5 samples and more,
// three workgroups per CU, where the clamped ring's extra loads cost more).
constexpr int PASS_PF_EX = 4;
template <bool EX, class PT, int MB, bool TR, int CB = 2>
__device__ __forceinline__ void pass_panel(const typename PT::T* As, int lda, const typename PT::T* Bp, int KBtot,
                                           int nb0, int nbstep, int ncb, int kb0, int kb1, f32x16_t (&acc)[CB][MB],
                                           int lane) {
  if constexpr (EX) {
    BRing<PT, CB, PASS_PF_EX> ring;
    panel_prefetch(ring, Bp, KBtot, nb0, nbstep, ncb, kb0, kb1, lane);
    panel_run<PT, MB, CB, PASS_PF_EX, TR>(ring, As, lda, Bp, KBtot, nb0, nbstep, ncb, kb0, kb1, 0, acc, lane);
  } else {
    mfma_panel<PT, MB, CB, PASS_PF, TR>(As, lda, Bp, KBtot, nb0, nbstep, ncb, kb0, kb1, 0, acc, lane);
  }
}

// EX launches (small grids) run 8 wavefronts per workgroup, each with half
// the column blocks and dim pairs of a 4-wavefront workgroup: twice the
// wavefronts per SIMD where the grid is too small to fill the CUs.
template <class PT, int MB, int KT, int NH, bool EX>
__global__ void __launch_bounds__(EX ? 512 : 256) __attribute__((amdgpu_waves_per_eu(3)))
flow_pass_kernel(FlowDev f, PassArgs a) {
  using T = typename PT::T;
  constexpr int MT = 32 * MB;
  constexpr int NW = EX ? 8 : 4, NT = 64 * NW;
  constexpr int CBH = NW == 4 ? 2 : 1;  // column blocks per hidden panel
  extern __shared__ __align__(16) unsigned char smem[];
  const FlowDims& dm = f.dm;
  PassSmem<PT> sm(dm, MT);
  float* xs = reinterpret_cast<float*>(smem + sm.xs);
  float* ldr = reinterpret_cast<float*>(smem + sm.ldr);
  float* ldp = reinterpret_cast<float*>(smem + sm.ldp);
  T* reg1 = reinterpret_cast<T*>(smem + sm.r1);
  T* reg2 = reinterpret_cast<T*>(smem + sm.r2);
  T* a0 = reg2;
  // hidden layer i lands in region 1 (i even) or 2 (i odd); a chunk's spline
  // parameters use the region the last hidden layer is not in
  T* hlast = ((NH - 1) & 1) ? reg2 : reg1;
  float* obuf = reinterpret_cast<float*>(((NH - 1) & 1) ? reg1 : reg2);
  // the wave index in a scalar register: branches on it are uniform
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int d = dm.d, ldx = sm.ldx;
  const int64_t m0 = (int64_t)blockIdx.x * MT;
  const float* th = f.theta;
  const T* pk = reinterpret_cast<const T*>(f.packed);
  // full-covariance base Gaussian: chol(base_cov) row-major, then sum log L_jj
  float* ub = reinterpret_cast<float*>(smem + sm.ub);
  const float* Lb = f.pbias + dm.pb_base();
  const bool dense_sample = dm.base_dense && a.dir != 0 && (a.flags & FL_SAMPLE);

  // ---- load tile (+ standardisation / base sampling)
  for (TileWalk tw(tid, d, NT); tw.m < MT; tw.next()) {
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
      } else if (a.dir != 0 && (a.flags & FL_SAMPLE) && !dense_sample) {
        v = th[dm.th_base_mean() + j] + sqrtf(th[dm.th_base_cov() + (int64_t)j * d + j]) * v;
      }
    }
    if (dense_sample) ub[m * ldx + j] = v;
    xs[m * ldx + j] = v;
  }
  __syncthreads();
  if (dense_sample) {  // base draw mean + L xi (jax.random.multivariate_normal, Cholesky method)
    for (TileWalk tw(tid, d, NT); tw.m < MT; tw.next()) {
      const int m = tw.m, j = tw.j;
      const float* Lr = Lb + (int64_t)j * d;
      float acc = th[dm.th_base_mean() + j];
      for (int k = 0; k <= j; ++k) acc += Lr[k] * ub[m * ldx + k];
      xs[m * ldx + j] = acc;
    }
    __syncthreads();
  }
  // exact inverse with BASE_LOGP: the pass returns the flow log-density of
  // its output, log N(z) - sum of the inverse log-dets; start from -log N(z)
  const bool inv_logq = (a.dir == 2) && (a.flags & FL_BASE);
  for (int m = tid; m < MT; m += NT) {
    float v = 0.f;
    if (inv_logq && dm.base_dense) {
      // -log N(z; mean, L L^T) = |u|^2 / 2 + sum log L_jj + d/2 log(2 pi), L u = z - mean
      float q = 0.f;
      if (dense_sample) {
        for (int j = 0; j < d; ++j) q += ub[m * ldx + j] * ub[m * ldx + j];
      } else {
        for (int j = 0; j < d; ++j) {
          const float* Lr = Lb + (int64_t)j * d;
          float r = xs[m * ldx + j] - th[dm.th_base_mean() + j];
          for (int k = 0; k < j; ++k) r -= Lr[k] * ub[m * ldx + k];
          const float u = r / Lr[j];
          ub[m * ldx + j] = u;
          q += u * u;
        }
      }
      v = 0.5f * q + Lb[(int64_t)d * d] + 0.91893853320467274f * (float)d;
    } else if (inv_logq) {
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
      for (TileWalk tw(tid, d, NT); tw.m < MT; tw.next()) {
        const int m = tw.m, j = tw.j;
        if (m0 + m < a.n) a.saved[((int64_t)l * a.n + m0 + m) * d + j] = xs[m * ldx + j];
      }
    }
    const float sc = th[dm.th_scale(l)], sh = th[dm.th_shift(l)];
    const bool affine_first = (a.dir != 2);
    // ---- ScalarAffine (forward / reference inverse apply it before the spline)
    if (affine_first) {
      const float e = expf(a.dir == 0 ? sc : -sc);
      for (TileWalk tw(tid, d, NT); tw.m < MT; tw.next()) {
        const int m = tw.m, j = tw.j;
        float v = xs[m * ldx + j];
        xs[m * ldx + j] = (a.dir == 0) ? (v + sh) * e : v * e - sh;
      }
      for (int m = tid; m < MT; m += NT) ldr[m] += (a.dir == 0 ? sc : -sc) * (float)d;
      __syncthreads();
    }
    // ---- conditioner input (conditioning dims only)
    const int dc = dm.dc(l), kc = dm.kc(l);
    for (TileWalk tw(tid, kc, NT); tw.m < MT; tw.next()) {
      const int m = tw.m, k = tw.j;
      a0[m * sm.la0 + k] = PT::cvt(k < dc ? xs[m * ldx + dm.cond(l, k)] : 0.f);
    }
    __syncthreads();
    const T* W = pk + dm.pk_layer(l);
    const float* bl = f.pbias + dm.pb_layer(l);
    const int hlp = dm.hlp();
    // ---- hidden layers: h_i = tanh(W_i h_{i-1} + b_i), h_{-1} = the
    // conditioning dims (W_0 restricted to them)
    {
      const T* in = a0;
      int ldin = sm.la0, KB = kc / PT::KSTEP;
#pragma unroll
      for (int i = 0; i < NH; ++i) {
        T* out = (i & 1) ? reg2 : reg1;
        const int NBt = dm.hp(i) / 32;
        for (int c0 = wave; c0 < NBt; c0 += CBH * NW) {
          const int ncb = (CBH == 2 && c0 + NW < NBt) ? 2 : 1;
          f32x16_t acc[CBH][MB];
          // the accumulators start from the bias (an unused second block from the first's)
#pragma unroll
          for (int j = 0; j < CBH; ++j) bias_tr<MB>(acc[j], bl + dm.pb_b(i), j < ncb ? c0 + NW * j : c0, lane);
          pass_panel<EX, PT, MB, true, CBH>(in, ldin, W + dm.pk_wh(l, i), KB, c0, NW, ncb, 0, KB, acc, lane);
#pragma unroll
          for (int j = 0; j < CBH; ++j)
            if (j < ncb) hidden_epilogue_tr<PT, MB>(out, sm.lh, nullptr, c0 + NW * j, acc[j], lane);
        }
        __syncthreads();
        in = out;
        ldin = sm.lh;
        KB = dm.hp(i) / PT::KSTEP;
      }
    }
    // ---- GEMM3 (transposed, register-spline order) + spline in registers
    if (dm.reg_spline()) {
      const int npair = dm.n_pairs(l), KB = hlp / PT::KSTEP, dtl = dm.dt(l), K = dm.K;
      const T* W3r = W + dm.pk_w3r(l);
      const float* b3r = bl + dm.pb_b3r(l);
      float* ldw = ldp;  // [NW waves][MT] partial log-dets
      float ldacc[MB];
#pragma unroll
      for (int r = 0; r < MB; ++r) ldacc[r] = 0.f;
      const int hf = lane >> 5;
      for (int q = wave; q < npair; q += NW) {
        const int k = 2 * q + hf;
        // the accumulators start from the 32 spline biases of the lane's
        // dim (a lane past the last dim reads the pair's first dim's)
        f32x16_t acc[2][MB];
        {
          const float4* bp = reinterpret_cast<const float4*>(b3r + q * 64 + (k < dtl ? hf * 32 : 0));
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const float4 b4 = bp[u];
#pragma unroll
            for (int r = 0; r < MB; ++r) {
              acc[u >> 2][r][4 * (u & 3)] = b4.x;
              acc[u >> 2][r][4 * (u & 3) + 1] = b4.y;
              acc[u >> 2][r][4 * (u & 3) + 2] = b4.z;
              acc[u >> 2][r][4 * (u & 3) + 3] = b4.w;
            }
          }
        }
        pass_panel<EX, PT, MB, true>(hlast, sm.lh, W3r, KB, 2 * q, 1, 2, 0, KB, acc, lane);
        if (k < dtl) {
          const int j = dm.trans(l, k);
#pragma unroll
          for (int r = 0; r < MB; ++r) {
            float pp[32];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
              pp[v] = acc[0][r][v];
              pp[16 + v] = acc[1][r][v];
            }
            const int m = r * 32 + (lane & 31);
            const float xv = xs[m * ldx + j];
            float yv, ldv;
            if (a.dir == 0)
              rs_fwd<KT, sizeof(T) == 2>(pp, K, f, xv, yv, ldv);
            else
              rs_inv<KT, sizeof(T) == 2>(pp, K, f, xv, yv, ldv);
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
      for (int m = tid; m < MT; m += NT) {
        float s = (ldw[m] + ldw[MT + m]) + (ldw[2 * MT + m] + ldw[3 * MT + m]);
        if constexpr (NW == 8) s += (ldw[4 * MT + m] + ldw[5 * MT + m]) + (ldw[6 * MT + m] + ldw[7 * MT + m]);
        ldr[m] += s;
      }
      __syncthreads();
    }
    // ---- GEMM3 in chunks of G transformed dims + spline epilogue (K > 10)
    const int nch = dm.reg_spline() ? 0 : dm.n_chunks(l), P = dm.P();
    for (int c = 0; c < nch; ++c) {
      const int nc = dm.chunk_nc(l, c), cnt = dm.chunk_cnt(l, c), coff = dm.chunk_off(l, c);
      const int KB = hlp / PT::KSTEP, NBt = nc / 32;
      for (int c0 = wave; c0 < NBt; c0 += CBH * NW) {
        const int ncb = (CBH == 2 && c0 + NW < NBt) ? 2 : 1;
        f32x16_t acc[CBH][MB] = {};
        mfma_panel<PT, MB, CBH, PASS_PF>(hlast, sm.lh, W + dm.pk_w3(l), KB, coff / 32 + c0, NW, ncb, 0, KB, 0, acc,
                                          lane);
#pragma unroll
        for (int j = 0; j < CBH; ++j) {
          if (j < ncb) {
            const int col = (c0 + NW * j) * 32 + (lane & 31);
            const float bb = bl[dm.pb_b3(l) + coff + col];
#pragma unroll
            for (int r = 0; r < MB; ++r)
#pragma unroll
              for (int v = 0; v < 16; ++v) obuf[(r * 32 + out_row(v, lane)) * sm.lo_ + col] = acc[j][r][v] + bb;
          }
        }
      }
      __syncthreads();
      for (int i = tid; i < MT * cnt; i += NT) {
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
      for (int m = tid; m < MT; m += NT) {
        float s = ldr[m];
        for (int jj = 0; jj < cnt; ++jj) s += ldp[m * dm.G() + jj];
        ldr[m] = s;
      }
      __syncthreads();
    }
    if (!affine_first) {  // exact inverse: affine^-1 after spline^-1
      const float e = expf(-sc);
      for (TileWalk tw(tid, d, NT); tw.m < MT; tw.next()) {
        const int m = tw.m, j = tw.j;
        xs[m * ldx + j] = xs[m * ldx + j] * e - sh;
      }
      for (int m = tid; m < MT; m += NT) ldr[m] += -sc * (float)d;
      __syncthreads();
    }
  }
  // ---- outputs
  for (int m = tid; m < MT; m += NT) {
    int64_t row = m0 + m;
    if (row >= a.n) continue;
    float ld = ldr[m];
    if (a.dir == 0 && (a.flags & FL_BASE) && dm.base_dense) {
      // log N(y; mean, L L^T) (common.py:285-286): forward substitution
      // L u = y - mean, then -|u|^2 / 2 - sum log L_jj - d/2 log(2 pi);
      // its y-gradient -Sigma^-1 (y - mean) = -L^-T u by back substitution
      float q = 0.f;
      for (int j = 0; j < d; ++j) {
        const float* Lr = Lb + (int64_t)j * d;
        float r = xs[m * ldx + j] - th[dm.th_base_mean() + j];
        for (int k = 0; k < j; ++k) r -= Lr[k] * ub[m * ldx + k];
        const float u = r / Lr[j];
        ub[m * ldx + j] = u;
        q += u * u;
      }
      if (a.gy_out) {
        for (int j = d - 1; j >= 0; --j) {
          float r = ub[m * ldx + j];
          for (int k = j + 1; k < d; ++k) r -= Lb[(int64_t)k * d + j] * ub[m * ldx + k];
          const float w = r / Lb[(int64_t)j * d + j];
          ub[m * ldx + j] = w;
          a.gy_out[row * d + j] = a.g_logp[row] * (-w);
        }
      }
      ld += -0.5f * q - Lb[(int64_t)d * d] - 0.91893853320467274f * (float)d;
    } else if (a.dir == 0 && (a.flags & FL_BASE)) {
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
    for (TileWalk tw(tid, d, NT); tw.m < MT; tw.next()) {
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

template <class PT, int KT, int NH>
static void launch_pass(dim3 g, dim3 b, size_t shm, hipStream_t st, const FlowDev& fd, const PassArgs& a) {
  if ((int)g.x <= 2 * device_cus()) {
    set_lds(flow_pass_kernel<PT, 1, KT, NH, true>, shm);
    hipLaunchKernelGGL((flow_pass_kernel<PT, 1, KT, NH, true>), g, dim3(512), shm, st, fd, a);
  } else {
    set_lds(flow_pass_kernel<PT, 1, KT, NH, false>, shm);
    hipLaunchKernelGGL((flow_pass_kernel<PT, 1, KT, NH, false>), g, b, shm, st, fd, a);
  }
}

template <class PT>
static void launch_pass_kt(int kt, int nh, dim3 g, dim3 b, size_t shm, hipStream_t st, const FlowDev& fd,
                           const PassArgs& a) {
#define FMC_PASS_NH(KT_)                                         \
  switch (nh) {                                                  \
    case 1: launch_pass<PT, KT_, 1>(g, b, shm, st, fd, a); break; \
    case 2: launch_pass<PT, KT_, 2>(g, b, shm, st, fd, a); break; \
    case 3: launch_pass<PT, KT_, 3>(g, b, shm, st, fd, a); break; \
    default: launch_pass<PT, KT_, 4>(g, b, shm, st, fd, a); break; \
  }
  if (kt == 8) {
    FMC_PASS_NH(8)
  } else if (kt == 10) {
    FMC_PASS_NH(10)
  } else {
    FMC_PASS_NH(0)
  }
#undef FMC_PASS_NH
}

}  // namespace fmc

using namespace fmc;

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
  const int kt = spline_kt(fd.dm.K);
  if (bf)
    launch_pass_kt<PBF16>(kt, fd.dm.nh, g, b, shm, st, fd, a);
  else
    launch_pass_kt<PF32>(kt, fd.dm.nh, g, b, shm, st, fd, a);
  return check_launch("flow pass");
}

