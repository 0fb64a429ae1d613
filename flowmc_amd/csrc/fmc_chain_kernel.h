// The fused chain kernel (one kernel per lane layout NV, step kind and target
// kind) and its host-side launch helpers: target-descriptor check, lane
// layout heuristic and launch. Shared by fmc_chains.hip (MALA, GRW, HMC with
// a unit or diagonal metric) and fmc_hmc_dense.hip (HMC with a dense
// condition matrix).
#pragma once

#include "fmc_chain_steps.h"

namespace fmc {

// KK_HMC runs the unit and diagonal condition matrices; a dense condition
// matrix has a kernel of its own (KK_HMC_DENSE, built in fmc_hmc_dense.hip):
// its momentum product would otherwise raise every HMC kernel's register
// allocation (and lower its occupancy).
enum KernelKind { KK_MALA = 0, KK_HMC = 1, KK_GRW = 2, KK_HMC_DENSE = 3 };

// One kernel per (lane layout NV, step kind, target kind TK): the target's
// value and gradient are inlined without a run-time dispatch. HMC keeps
// position, momentum, gradient, metric and the leapfrog copies in registers:
// its workgroups are bounded at 512 threads, which gives it 256 VGPRs
// instead of the 128 of a 1024-thread bound (spills at NV >= 2).
template <int NV, int KIND, int TK>
__global__ void __launch_bounds__((KIND == KK_HMC || KIND == KK_HMC_DENSE) ? 512 : 1024) chain_kernel(fmc_target tg, ChainIO io, StepCfg cfg,
                                                     int lpc, int n_lds) {
  extern __shared__ float smem[];
  TargetDev t = stage_target(tg, smem, n_lds);
  const int lane = threadIdx.x & 63;
  const int r = lane & (lpc - 1);
  const int gbase = lane - r;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x / lpc) + threadIdx.x / lpc;
  if (c >= io.C) return;  // whole group leaves together
  if (KIND == KK_MALA)
    mala_group<NV, TK>(t, io, cfg, c, r, lpc, gbase);
  else if (KIND == KK_GRW)
    grw_group<NV, TK>(t, io, cfg, c, r, lpc, gbase);
  else if (KIND == KK_HMC_DENSE)
    hmc_group<NV, MK_DENSE, TK>(t, io, cfg, c, r, lpc, gbase);
  else if (io.metric)
    hmc_group<NV, MK_DIAG, TK>(t, io, cfg, c, r, lpc, gbase);
  else
    hmc_group<NV, MK_UNIT, TK>(t, io, cfg, c, r, lpc, gbase);
}

// ----------------------------------------------------------- host side ----
static inline int check_target(const fmc_target* t) {
  FMC_ARG(t != nullptr, "target is NULL");
  FMC_ARG(t->n_dim >= 1, "target n_dim must be >= 1 (got %d)", t->n_dim);
  FMC_ARG(t->kind >= FMC_TARGET_GAUSS_DIAG && t->kind <= FMC_TARGET_DUAL_MOON, "unknown target kind %d",
          t->kind);
  if (t->kind == FMC_TARGET_GMM_DIAG) FMC_ARG(t->n_modes >= 1, "GMM target needs n_modes >= 1");
  if (t->kind == FMC_TARGET_DUAL_MOON) FMC_ARG(t->n_dim >= 2, "dual-moon target needs n_dim >= 2");
  int64_t need = target_param_count(t->kind, t->n_dim, t->n_modes);
  FMC_ARG(t->n_params >= need, "target kind %d with d=%d needs %lld params, got %lld", t->kind, t->n_dim,
          (long long)need, (long long)t->n_params);
  FMC_ARG(need == 0 || t->params != nullptr, "target params pointer is NULL");
  return FMC_OK;
}

struct Layout {
  int lpc, nv, block;
};

static inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Heuristic lane layout: aim for ~4 dims per lane for long vectors (one float4
// chunk, so the fused step's register footprint stays small), but keep whole
// chains in one lane when d is tiny. Rosenbrock takes two chunks per lane: its
// three neighbour reads per chunk and the group sum amortise over twice the
// elements (d = 128: 16 x 2 beats 32 x 1 by 6 % for MALA and 7 % for GRW),
// while the mixture's per-mode loop is 8 % slower at two chunks per lane.
static inline Layout choose_layout(int d, int kind, const fmc_launch_cfg* cfg) {
  Layout L{0, 0, 256};
  if (cfg && cfg->lanes_per_chain > 0 && cfg->vec_per_lane > 0) {
    L.lpc = cfg->lanes_per_chain;
    L.nv = cfg->vec_per_lane;
  } else {
    int chunks = (d + 3) / 4;
    if (kind == FMC_TARGET_GAUSS_DENSE) {
      // O(d^2) gather: keep the chain in few lanes.
      L.lpc = 1;
      while (L.lpc * 8 < chunks) L.lpc *= 2;
      L.nv = (chunks + L.lpc - 1) / L.lpc;
    } else {
      const int per_lane = (kind == FMC_TARGET_ROSENBROCK) ? 2 : 1;
      L.lpc = 1;
      while (L.lpc * per_lane < chunks && L.lpc < 64) L.lpc *= 2;
      L.nv = (chunks + L.lpc - 1) / L.lpc;
    }
    if (L.nv > 2) L.nv = (L.nv <= 4) ? 4 : 8;
  }
  if (cfg && cfg->block_size > 0) L.block = cfg->block_size;
  return L;
}

static inline int check_layout(const Layout& L, int d) {
  FMC_ARG(is_pow2(L.lpc) && L.lpc <= 64, "lanes_per_chain must be a power of two <= 64 (got %d)", L.lpc);
  FMC_ARG(L.nv == 1 || L.nv == 2 || L.nv == 4 || L.nv == 8, "vec_per_lane must be 1, 2, 4 or 8 (got %d)",
          L.nv);
  FMC_ARG(4LL * L.nv * L.lpc >= d, "layout lanes_per_chain=%d vec_per_lane=%d holds %d dims < d=%d", L.lpc,
          L.nv, 4 * L.nv * L.lpc, d);
  FMC_ARG(L.block >= 64 && L.block <= 1024 && L.block % 64 == 0, "block_size must be a multiple of 64 in [64,1024]");
  return FMC_OK;
}

static inline int lds_floats(const fmc_target* t) {
  int64_t n = target_param_count(t->kind, t->n_dim, t->n_modes);
  return (n > 0 && n * 4 <= 48 * 1024) ? (int)n : 0;
}

template <int KIND, int TK>
static void launch_chain_nv(hipStream_t st, const fmc_target* t, const ChainIO& io, const StepCfg& cfg,
                            const Layout& L, dim3 g, dim3 b, size_t shm, int n_lds) {
  switch (L.nv) {
    case 1: hipLaunchKernelGGL((chain_kernel<1, KIND, TK>), g, b, shm, st, *t, io, cfg, L.lpc, n_lds); break;
    case 2: hipLaunchKernelGGL((chain_kernel<2, KIND, TK>), g, b, shm, st, *t, io, cfg, L.lpc, n_lds); break;
    case 4: hipLaunchKernelGGL((chain_kernel<4, KIND, TK>), g, b, shm, st, *t, io, cfg, L.lpc, n_lds); break;
    default: hipLaunchKernelGGL((chain_kernel<8, KIND, TK>), g, b, shm, st, *t, io, cfg, L.lpc, n_lds); break;
  }
}

template <int KIND>
static int launch_chain_impl(hipStream_t st, const fmc_target* t, const ChainIO& io, const StepCfg& cfg,
                             const Layout& L) {
  FMC_ARG((KIND != KK_HMC && KIND != KK_HMC_DENSE) || L.block <= 512,
          "fused HMC launches take at most 512 threads per workgroup (got %d)", L.block);
  const int per_block = L.block / L.lpc;
  const int64_t grid = (io.C + per_block - 1) / per_block;
  FMC_ARG(grid <= 0x7fffffffLL, "too many chains for one launch");
  const int n_lds = lds_floats(t);
  const size_t shm = (size_t)n_lds * sizeof(float);
  dim3 g((unsigned)grid), b((unsigned)L.block);
  switch (t->kind) {  // check_target has accepted the kind
    case FMC_TARGET_GAUSS_DIAG: launch_chain_nv<KIND, FMC_TARGET_GAUSS_DIAG>(st, t, io, cfg, L, g, b, shm, n_lds); break;
    case FMC_TARGET_GAUSS_DENSE: launch_chain_nv<KIND, FMC_TARGET_GAUSS_DENSE>(st, t, io, cfg, L, g, b, shm, n_lds); break;
    case FMC_TARGET_ROSENBROCK: launch_chain_nv<KIND, FMC_TARGET_ROSENBROCK>(st, t, io, cfg, L, g, b, shm, n_lds); break;
    case FMC_TARGET_FUNNEL: launch_chain_nv<KIND, FMC_TARGET_FUNNEL>(st, t, io, cfg, L, g, b, shm, n_lds); break;
    case FMC_TARGET_GMM_DIAG: launch_chain_nv<KIND, FMC_TARGET_GMM_DIAG>(st, t, io, cfg, L, g, b, shm, n_lds); break;
    default: launch_chain_nv<KIND, FMC_TARGET_DUAL_MOON>(st, t, io, cfg, L, g, b, shm, n_lds); break;
  }
  return check_launch(KIND == KK_MALA ? "mala chain kernel" : KIND == KK_GRW ? "grw chain kernel" : "hmc chain kernel");
}

// Fused HMC with a dense condition matrix (defined in fmc_hmc_dense.hip).
int launch_hmc_dense(hipStream_t st, const fmc_target* t, const ChainIO& io, const StepCfg& cfg, const Layout& L);

template <int KIND>
static int launch_chain(hipStream_t st, const fmc_target* t, const ChainIO& io, const StepCfg& cfg,
                        const Layout& L) {
  if constexpr (KIND == KK_HMC) {
    if (io.chol) return launch_hmc_dense(st, t, io, cfg, L);
  }
  return launch_chain_impl<KIND>(st, t, io, cfg, L);
}

}  // namespace fmc
