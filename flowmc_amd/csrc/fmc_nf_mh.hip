// Independence Metropolis-Hastings over normalizing-flow proposals
// (NFProposal.kernel's scan, src/flowMC/resource/kernel/NF_proposal.py:91-126)
// and the two C-ABI entry points that run it: fmc_nf_mh, whose caller passes
// the [C, S] proposal-index buffer, and fmc_nf_mh_noalloc, which keeps that
// buffer in a fixed device array of this library instead.
#include <mutex>

#include "fmc_common.h"

namespace fmc {

// --------------------------------------------------------- NF MH scan -----
// NF_proposal.py:91-126: sequential independence MH over the S proposals of
// each chain. Phase 1 (one thread per chain) decides and records which
// proposal the chain sits on after every step; phase 2 gathers the rows.
__global__ void nf_mh_scan_kernel(const float* __restrict__ lp0, const float* __restrict__ q0,
                                  const float* __restrict__ lpp, const float* __restrict__ qp,
                                  int64_t C, int S, uint32_t k0, uint32_t k1, const float* log_u,
                                  int* out_idx, float* lp_last, float* out_logp, float* out_acc,
                                  int64_t buf_len, int64_t cursor, int thinning) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float lpc = lp0[c], qc = q0[c], lp_keep = lp0[c];
  int idx = -1;
  const int s_last = ((S - 1) / thinning) * thinning;
  for (int s = 0; s < S; ++s) {
    const int64_t k = c * S + s;
    float lpv = lpp[k], qv = qp[k];
    float ratio = (lpv - lpc) - (qv - qc);
    float lu = log_u ? log_u[k] : philox_log_uniform(k0, k1, (uint32_t)c, (uint32_t)s, RNG_NF_UNIFORM);
    bool acc = lu < ratio;
    if (acc) {
      idx = s;
      lpc = lpv;
      qc = qv;
    }
    out_idx[k] = idx;
    if (s == s_last) lp_keep = lpc;
    if (s % thinning == 0) {
      int64_t row = cursor + s / thinning;
      if (out_logp) out_logp[c * buf_len + row] = lpc;
      if (out_acc) out_acc[c * buf_len + row] = acc ? 1.f : 0.f;
    }
  }
  if (lp_last) lp_last[c] = lp_keep;
}

// One (chain, thinned step) row per 64 lanes.
__global__ void nf_mh_gather_kernel(const float* __restrict__ x0, const float* __restrict__ prop,
                                    const int* __restrict__ idx, int64_t C, int S, int d,
                                    float* out_pos, int64_t buf_len, int64_t cursor, int thinning,
                                    int n_rows) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= C * n_rows) return;
  const int64_t c = w / n_rows;
  const int rr = (int)(w - c * n_rows);
  const int s = rr * thinning;
  const int id = idx[c * S + s];
  const float* src = (id < 0) ? x0 + c * d : prop + (c * S + id) * d;
  float* dst = out_pos + (c * buf_len + cursor + rr) * d;
  for (int j = lane; j < d; j += 64) dst[j] = src[j];
}

// Final position of each chain after the scan.
__global__ void nf_mh_last_kernel(const float* __restrict__ x0, const float* __restrict__ prop,
                                  const int* __restrict__ idx, int64_t C, int S, int d, int s_last,
                                  float* x_last) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C * d) return;
  const int64_t c = i / d;
  const int j = (int)(i - c * d);
  const int id = idx[c * S + s_last];
  x_last[i] = (id < 0) ? x0[i] : prop[(c * S + id) * d + j];
}

// Per-(chain, step) proposal indices of fmc_nf_mh_noalloc: a fixed array of
// FMC_NF_MH_NOALLOC_MAX int32 (16 MiB) on every device the library runs on,
// held until the process ends. The scan writes every entry it later reads,
// so nothing needs clearing between calls; calls that overlap on one device
// share it and race.
static __device__ __align__(256) int32_t scratch[FMC_NF_MH_NOALLOC_MAX];

// Device address of `scratch` on the current device, taken once per device
// with hipGetSymbolAddress (the array's name is not a device address in host
// code).
static int scratch_address(int32_t** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static int32_t* addr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  FMC_REQUIRE(e == hipSuccess, FMC_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  FMC_REQUIRE(dev >= 0 && dev < kMaxDevices, FMC_ERR_UNSUPPORTED, "device %d beyond the %d this library tracks", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    FMC_REQUIRE(e == hipSuccess && p != nullptr, FMC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                hipGetErrorString(e));
    addr[dev] = static_cast<int32_t*>(p);
  }
  *out = addr[dev];
  return FMC_OK;
}

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Argument checks and the scan / gather / last-row launches shared by both
// entry points; idx is the [C, S] proposal-index buffer.
static int nf_mh_run(hipStream_t st, const float* x0, const float* lp0, const float* q0, const float* prop,
                     const float* lp_prop, const float* q_prop, int64_t n_chains, int32_t n_steps, int32_t n_dim,
                     uint32_t key0, uint32_t key1, const float* log_u, int32_t* idx, float* x_last, float* lp_last,
                     float* out_pos, float* out_logp, float* out_acc, int64_t buf_len, int64_t cursor,
                     int32_t thinning) {
  const int n_rows = (n_steps + thinning - 1) / thinning;
  hipLaunchKernelGGL(nf_mh_scan_kernel, dim3(grid1(n_chains, 256)), dim3(256), 0, st, lp0, q0, lp_prop, q_prop,
                     n_chains, n_steps, key0, key1, log_u, idx, lp_last, out_logp, out_acc, buf_len, cursor,
                     thinning);
  int rc;
  if ((rc = check_launch("nf mh scan"))) return rc;
  if (out_pos) {
    hipLaunchKernelGGL(nf_mh_gather_kernel, dim3(grid1(n_chains * n_rows, 4)), dim3(256), 0, st, x0, prop, idx,
                       n_chains, n_steps, n_dim, out_pos, buf_len, cursor, thinning, n_rows);
    if ((rc = check_launch("nf mh gather"))) return rc;
  }
  if (x_last) {
    const int s_last = (n_rows - 1) * thinning;
    hipLaunchKernelGGL(nf_mh_last_kernel, dim3(grid1(n_chains * n_dim, 256)), dim3(256), 0, st, x0, prop, idx,
                       n_chains, n_steps, n_dim, s_last, x_last);
    if ((rc = check_launch("nf mh last"))) return rc;
  }
  return FMC_OK;
}

static int nf_mh_check(const float* x0, const float* lp0, const float* q0, const float* prop, const float* lp_prop,
                       const float* q_prop, int64_t n_chains, int32_t n_steps, int32_t n_dim, const float* out_pos,
                       const float* out_logp, const float* out_acc, int64_t buf_len, int64_t cursor,
                       int32_t thinning) {
  FMC_ARG(x0 && lp0 && q0 && prop && lp_prop && q_prop, "NULL pointer");
  FMC_ARG(n_chains >= 0 && n_steps >= 1 && n_dim >= 1, "bad sizes");
  return check_buffers(out_pos, out_logp, out_acc, buf_len, cursor, thinning, 0, n_steps);
}

}  // namespace fmc

using namespace fmc;

extern "C" int fmc_nf_mh(void* stream, const float* x0, const float* lp0, const float* q0, const float* prop,
                         const float* lp_prop, const float* q_prop, int64_t n_chains, int32_t n_steps,
                         int32_t n_dim, uint32_t key0, uint32_t key1, const float* log_u, int32_t* out_idx,
                         float* x_last, float* lp_last, float* out_pos, float* out_logp, float* out_acc,
                         int64_t buf_len, int64_t cursor, int32_t thinning) {
  int rc = nf_mh_check(x0, lp0, q0, prop, lp_prop, q_prop, n_chains, n_steps, n_dim, out_pos, out_logp, out_acc,
                       buf_len, cursor, thinning);
  if (rc) return rc;
  FMC_ARG(out_idx != nullptr, "NULL pointer");
  if (n_chains == 0) return FMC_OK;
  return nf_mh_run((hipStream_t)stream, x0, lp0, q0, prop, lp_prop, q_prop, n_chains, n_steps, n_dim, key0, key1,
                   log_u, out_idx, x_last, lp_last, out_pos, out_logp, out_acc, buf_len, cursor, thinning);
}

extern "C" int fmc_nf_mh_noalloc(void* stream, const float* x0, const float* lp0, const float* q0,
                                 const float* prop, const float* lp_prop, const float* q_prop, int64_t n_chains,
                                 int32_t n_steps, int32_t n_dim, uint32_t key0, uint32_t key1, const float* log_u,
                                 float* x_last, float* lp_last, float* out_pos, float* out_logp, float* out_acc,
                                 int64_t buf_len, int64_t cursor, int32_t thinning) {
  int rc = nf_mh_check(x0, lp0, q0, prop, lp_prop, q_prop, n_chains, n_steps, n_dim, out_pos, out_logp, out_acc,
                       buf_len, cursor, thinning);
  if (rc) return rc;
  FMC_ARG(n_chains <= (int64_t)FMC_NF_MH_NOALLOC_MAX / n_steps,
          "n_chains * n_steps = %lld exceeds the %d proposal slots of the no-allocation scan",
          (long long)n_chains * n_steps, (int)FMC_NF_MH_NOALLOC_MAX);
  int32_t* idx = nullptr;
  if ((rc = scratch_address(&idx))) return rc;
  if (n_chains == 0) return FMC_OK;
  return nf_mh_run((hipStream_t)stream, x0, lp0, q0, prop, lp_prop, q_prop, n_chains, n_steps, n_dim, key0, key1,
                   log_u, idx, x_last, lp_last, out_pos, out_logp, out_acc, buf_len, cursor, thinning);
}
