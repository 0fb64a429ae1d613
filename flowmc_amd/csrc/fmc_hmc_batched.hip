// Many HMC problems of uneven dimension advanced in one launch (gfx950).
//
// Each work item (slot, first chain, n chains) is one workgroup; a chain is
// held by the slot's lanes_per_chain lanes exactly as in the single-target
// fused HMC kernel, and runs HMC.kernel (src/flowMC/resource/kernel/HMC.py:
// 98-151) for cfg.n_steps steps. Slots may mix target kinds, so the target
// is dispatched at run time (a workgroup-uniform branch). Specialising each
// workgroup's whole HMC loop on its target kind instead (one loop body per
// kind in one kernel) raised the register demand to the 128-VGPR cap with
// spills at two float4 chunks per lane, so the single loop with the uniform
// dispatch is kept. Its own translation unit: it builds in parallel with the
// other chain kernels.
#include <type_traits>

#include "fmc_chain_steps.h"

namespace fmc {

// At most 512 threads per workgroup: the register budget is then 256 VGPRs,
// which the run-time target dispatch needs at two or more float4 chunks per
// lane (under a 1024-thread bound the 128-VGPR cap spilled). The batch has
// about 1.4 wavefronts per SIMD, so the bound costs no occupancy.
template <int NV>
__global__ void __launch_bounds__(512) hmc_batched_kernel(const fmc_hmc_slot* __restrict__ slots,
                                                           const fmc_work_item* __restrict__ items,
                                                           StepCfg cfg) {
  const fmc_work_item it = items[blockIdx.x];
  const fmc_hmc_slot sl = slots[it.slot];
  const int lpc = sl.lanes_per_chain;
  const int lane = threadIdx.x & 63;
  const int r = lane & (lpc - 1);
  const int gbase = lane - r;
  const int local = threadIdx.x / lpc;
  if (local >= it.n_chains) return;  // whole chain groups leave together
  const int64_t c = (int64_t)it.chain0 + local;
  TargetDev t{sl.target.kind, sl.target.n_dim, sl.target.n_modes, sl.target.params};
  ChainIO io{sl.x,        sl.logp,     sl.metric,  sl.noise,  sl.log_u,  sl.out_pos, sl.out_logp,
             sl.out_acc,  sl.n_chains, sl.buf_len, sl.cursor, sl.key0,   sl.key1};
  io.chol = sl.momentum_chol;
  if (io.chol)
    hmc_group<NV, MK_DENSE, -1>(t, io, cfg, c, r, lpc, gbase);
  else if (io.metric)
    hmc_group<NV, MK_DIAG, -1>(t, io, cfg, c, r, lpc, gbase);
  else
    hmc_group<NV, MK_UNIT, -1>(t, io, cfg, c, r, lpc, gbase);
}

// Homogeneous unit-metric batches (every slot holds the same target kind TK
// and no condition matrix, the default HMC of HMC.py:25-40): the target is
// fixed at compile time and the slot's lanes_per_chain is turned into a
// compile-time constant by a workgroup-uniform switch (one copy of the step
// loop per power of two), so the group sums and broadcasts of every gradient
// evaluation are straight-line DPP code with no run-time branches. A slot
// that breaks the promise (another kind, a metric) is not advanced and its
// log-prob state is set to NaN.
template <int NV, int TK, int LPC>
__device__ __forceinline__ void hmc_unit_fixed(const fmc_work_item& it, const fmc_hmc_slot& sl,
                                               const StepCfg& cfg) {
  const int lane = threadIdx.x & 63;
  const int r = lane & (LPC - 1);
  const int gbase = lane - r;
  const int local = threadIdx.x / LPC;
  if (local >= it.n_chains) return;
  const int64_t c = (int64_t)it.chain0 + local;
  TargetDev t{TK, sl.target.n_dim, sl.target.n_modes, sl.target.params};
  ChainIO io{sl.x,        sl.logp,     nullptr,    sl.noise,  sl.log_u,  sl.out_pos, sl.out_logp,
             sl.out_acc,  sl.n_chains, sl.buf_len, sl.cursor, sl.key0,   sl.key1};
  hmc_group<NV, MK_UNIT, TK>(t, io, cfg, c, r, LPC, gbase);
}

template <int NV, int TK>
__global__ void __launch_bounds__(512) hmc_batched_unit_kernel(const fmc_hmc_slot* __restrict__ slots,
                                                                const fmc_work_item* __restrict__ items,
                                                                StepCfg cfg) {
  const fmc_work_item it = items[blockIdx.x];
  const fmc_hmc_slot sl = slots[it.slot];
  if (sl.target.kind != TK || sl.metric != nullptr || sl.momentum_chol != nullptr) {
    for (int i = threadIdx.x; i < it.n_chains; i += blockDim.x)
      if (sl.logp) sl.logp[it.chain0 + i] = __builtin_nanf("");
    return;
  }
  switch (sl.lanes_per_chain) {
    case 1: hmc_unit_fixed<NV, TK, 1>(it, sl, cfg); break;
    case 2: hmc_unit_fixed<NV, TK, 2>(it, sl, cfg); break;
    case 4: hmc_unit_fixed<NV, TK, 4>(it, sl, cfg); break;
    case 8: hmc_unit_fixed<NV, TK, 8>(it, sl, cfg); break;
    case 16: hmc_unit_fixed<NV, TK, 16>(it, sl, cfg); break;
    case 32: hmc_unit_fixed<NV, TK, 32>(it, sl, cfg); break;
    default: hmc_unit_fixed<NV, TK, 64>(it, sl, cfg); break;
  }
}

}  // namespace fmc

using namespace fmc;

extern "C" int fmc_hmc_run_batched(void* stream, const fmc_hmc_slot* slots, int32_t n_slots,
                                   const fmc_work_item* items, int32_t n_items, int32_t max_dim,
                                   int32_t n_steps, int32_t step0, float step_size, int32_t n_leapfrog,
                                   int32_t thinning, int32_t vec_per_lane, int32_t block_size,
                                   int32_t uniform_kind) {
  FMC_ARG(slots != nullptr && items != nullptr, "slots/items are NULL");
  FMC_ARG(n_slots >= 1 && n_items >= 0, "bad n_slots / n_items");
  FMC_ARG(n_leapfrog >= 0 && n_steps >= 0 && step0 >= 0 && thinning >= 1, "bad step arguments");
  FMC_ARG(vec_per_lane == 1 || vec_per_lane == 2 || vec_per_lane == 4 || vec_per_lane == 8,
          "vec_per_lane must be 1, 2, 4 or 8");
  FMC_ARG(block_size >= 64 && block_size <= 512 && block_size % 64 == 0,
          "block_size must be a multiple of 64 in [64, 512] (got %d)", block_size);
  FMC_ARG(max_dim >= 1, "max_dim must be >= 1");
  FMC_ARG(uniform_kind >= -1 && uniform_kind <= FMC_TARGET_DUAL_MOON,
          "uniform_kind must be -1 or a target kind (got %d)", uniform_kind);
  // The slot / item tables live in device memory; the host-side builder
  // (flowmc_amd.resource.kernel.batched_HMC.BatchedHMC) validates
  // lanes_per_chain (a power of two <= 64 dividing the block), dims and tiles.
  if (n_items == 0) return FMC_OK;
  StepCfg sc{n_steps, step0, step_size, n_leapfrog, thinning};
  hipStream_t st = (hipStream_t)stream;
  dim3 g((unsigned)n_items), b((unsigned)block_size);
  if (uniform_kind >= 0 && vec_per_lane <= 2) {
    // every slot: target kind uniform_kind, unit metric (the caller's promise)
    auto pick = [&](auto nv) {
      constexpr int NV = decltype(nv)::value;
      switch (uniform_kind) {
        case FMC_TARGET_GAUSS_DIAG: return hmc_batched_unit_kernel<NV, FMC_TARGET_GAUSS_DIAG>;
        case FMC_TARGET_GAUSS_DENSE: return hmc_batched_unit_kernel<NV, FMC_TARGET_GAUSS_DENSE>;
        case FMC_TARGET_ROSENBROCK: return hmc_batched_unit_kernel<NV, FMC_TARGET_ROSENBROCK>;
        case FMC_TARGET_FUNNEL: return hmc_batched_unit_kernel<NV, FMC_TARGET_FUNNEL>;
        case FMC_TARGET_GMM_DIAG: return hmc_batched_unit_kernel<NV, FMC_TARGET_GMM_DIAG>;
        default: return hmc_batched_unit_kernel<NV, FMC_TARGET_DUAL_MOON>;
      }
    };
    auto k = vec_per_lane == 1 ? pick(std::integral_constant<int, 1>{}) : pick(std::integral_constant<int, 2>{});
    hipLaunchKernelGGL(k, g, b, 0, st, slots, items, sc);
    return check_launch("batched hmc kernel (uniform unit-metric)");
  }
  switch (vec_per_lane) {
    case 1: hipLaunchKernelGGL(hmc_batched_kernel<1>, g, b, 0, st, slots, items, sc); break;
    case 2: hipLaunchKernelGGL(hmc_batched_kernel<2>, g, b, 0, st, slots, items, sc); break;
    case 4: hipLaunchKernelGGL(hmc_batched_kernel<4>, g, b, 0, st, slots, items, sc); break;
    default: hipLaunchKernelGGL(hmc_batched_kernel<8>, g, b, 0, st, slots, items, sc); break;
  }
  return check_launch("batched hmc kernel");
}
