// Fused HMC with a dense condition matrix M (HMC.kernel,
// src/flowMC/resource/kernel/HMC.py:98-151 with the momentum
// xi @ chol(inv(M)).T of :133-136 and the broadcast kinetic of :118-121):
// the chain kernels of fmc_chain_kernel.h instantiated for that step kind,
// in a translation unit of their own.
#include "fmc_chain_kernel.h"

namespace fmc {

int launch_hmc_dense(hipStream_t st, const fmc_target* t, const ChainIO& io, const StepCfg& cfg, const Layout& L) {
  return launch_chain_impl<KK_HMC_DENSE>(st, t, io, cfg, L);
}

}  // namespace fmc
