// Shared device/host helpers for the flowMC gfx950 kernels:
// error reporting, Philox4x32-10, Box–Muller normals, lane-group reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/flowmc.h"

namespace fmc {

// ---------------------------------------------------------------- errors ---
void set_error(int code, const char* fmt, ...);
// Converts a pending launch error into a status (and message).
int check_launch(const char* what);

#define FMC_REQUIRE(cond, code, ...)        \
  do {                                      \
    if (!(cond)) {                          \
      ::fmc::set_error((code), __VA_ARGS__); \
      return (code);                        \
    }                                       \
  } while (0)

#define FMC_ARG(cond, ...) FMC_REQUIRE(cond, FMC_ERR_INVALID_ARG, __VA_ARGS__)

// ---------------------------------------------------------------- Philox ---
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with the published
// multipliers and Weyl key increments.
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi,
                                                        uint32_t& lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  hi = __umulhi(a, b);
  lo = a * b;
#else
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
#endif
}

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c.x, hi0, lo0);
    philox_mulhilo(M1, c.z, hi1, lo1);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Uniform in the open interval (0,1): 23 random bits + 1/2 ulp offset, exact in FP32.
__host__ __device__ __forceinline__ float u01(uint32_t b) {
  return ((float)(b >> 9) + 0.5f) * 1.1920928955078125e-07f;  // 2^-23
}

// Stream ids of the counter's 4th word (one per use of randomness).
enum RngStream : uint32_t {
  RNG_MALA_NOISE = 0,
  RNG_MALA_UNIFORM = 1,
  RNG_HMC_MOMENTUM = 2,
  RNG_HMC_UNIFORM = 3,
  RNG_NF_UNIFORM = 4,
  RNG_FLOW_BASE = 5,
  RNG_GRW_NOISE = 7,
  RNG_GRW_UNIFORM = 8,
  RNG_PT_SWAP = 9,
};

// Four standard normals for dims 4*chunk .. 4*chunk+3 of (chain, step):
// Box–Muller on the pairs (w0,w1) and (w2,w3).
__device__ __forceinline__ float4 philox_normal4(uint32_t k0, uint32_t k1, uint32_t chunk,
                                                 uint32_t chain, uint32_t step, uint32_t sid) {
  u32x4 b = philox4x32_10(u32x4{chunk, chain, step, sid}, k0, k1);
  float r0 = sqrtf(-2.0f * logf(u01(b.x)));
  float r1 = sqrtf(-2.0f * logf(u01(b.z)));
  float s0, c0, s1, c1;
  sincospif(2.0f * u01(b.y), &s0, &c0);
  sincospif(2.0f * u01(b.w), &s1, &c1);
  return make_float4(r0 * c0, r0 * s0, r1 * c1, r1 * s1);
}

__device__ __forceinline__ float philox_log_uniform(uint32_t k0, uint32_t k1, uint32_t chain,
                                                    uint32_t step, uint32_t sid) {
  u32x4 b = philox4x32_10(u32x4{0u, chain, step, sid}, k0, k1);
  return logf(u01(b.x));
}

// ------------------------------------------------------- lane-group sums ---
// A chain owns `lpc` consecutive lanes (power of two, aligned); a butterfly
// over xor offsets < lpc stays inside the group and leaves the bitwise same
// sum in every lane of the group.
__device__ __forceinline__ float group_sum(float v, int lpc) {
  for (int o = lpc >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double group_sum(double v, int lpc) {
  for (int o = lpc >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float group_max(float v, int lpc) {
  for (int o = lpc >> 1; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ float& f4(float4& v, int e) { return (&v.x)[e]; }
__device__ __forceinline__ float f4c(const float4& v, int e) { return (&v.x)[e]; }

}  // namespace fmc
