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

// Row range of the thinned buffer writes of steps step0 .. step0+n_steps-1:
// global step t with t % thinning == 0 lands on row cursor + t / thinning.
static inline int check_buffers(const float* out_pos, const float* out_logp, const float* out_acc,
                                int64_t buf_len, int64_t cursor, int thinning, int step0, int n_steps) {
  FMC_ARG(thinning >= 1, "thinning must be >= 1");
  FMC_ARG(step0 >= 0 && n_steps >= 0, "step0 and n_steps must be >= 0");
  if (out_pos || out_logp || out_acc) {
    FMC_ARG(buf_len >= 1 && cursor >= 0, "bad buffer length / cursor");
    if (n_steps > 0) {
      int64_t last_t = step0 + n_steps - 1;
      int64_t last_row = cursor + last_t / thinning;
      FMC_ARG(last_row < buf_len, "buffer row %lld out of range (buf_len %lld)", (long long)last_row,
              (long long)buf_len);
    }
  }
  return FMC_OK;
}

// ---------------------------------------------------------------- Philox ---
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with the published
// multipliers and Weyl key increments.
struct u32x4 {
  uint32_t x, y, z, w;
};

// One 32x32 -> 64-bit product (a single v_mad_u64_u32 on the GPU instead of
// v_mul_hi_u32 + v_mul_lo_u32).
__host__ __device__ __forceinline__ void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi,
                                                        uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
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
// Box–Muller on the pairs (w0,w1) and (w2,w3), r = sqrt(-2 ln u0) and angle
// 2 pi u1. The transcendentals are the hardware ones: v_log_f32 (log2; u is
// in [2^-24, 1), no denormals), v_sqrt_f32, and v_sin_f32 / v_cos_f32, whose
// argument is in revolutions, so sin(2 pi u) needs no range reduction.
// Measured against FP64 over all 2^23 uniforms: log and r within 1.7e-7
// relative, sin / cos within 1.3e-7 absolute (the library forms: 1.9e-7,
// 1.5e-7, 5e-8), i.e. each normal within ~1.5e-6 absolute, at 101 instead
// of 229 VALU instructions per four normals.
__device__ __forceinline__ float4 philox_normal4(uint32_t k0, uint32_t k1, uint32_t chunk,
                                                 uint32_t chain, uint32_t step, uint32_t sid) {
  u32x4 b = philox4x32_10(u32x4{chunk, chain, step, sid}, k0, k1);
  const float m2ln2 = -1.38629436111989061f;  // -2 ln 2
  const float r0 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u01(b.x)));
  const float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u01(b.z)));
  const float a0 = u01(b.y), a1 = u01(b.w);
  return make_float4(r0 * __builtin_amdgcn_cosf(a0), r0 * __builtin_amdgcn_sinf(a0),
                     r1 * __builtin_amdgcn_cosf(a1), r1 * __builtin_amdgcn_sinf(a1));
}

// log u of (chain, step): hardware log2 times ln 2 (u >= 2^-24: no denormals).
__device__ __forceinline__ float philox_log_uniform(uint32_t k0, uint32_t k1, uint32_t chain,
                                                    uint32_t step, uint32_t sid) {
  u32x4 b = philox4x32_10(u32x4{0u, chain, step, sid}, k0, k1);
  return __builtin_amdgcn_logf(u01(b.x)) * 0.693147180559945309f;
}

// ------------------------------------------------------- lane-group sums ---
// A chain owns `lpc` consecutive lanes (power of two, aligned, the same for
// the whole wave). The collectives below move data on the VALU, not through
// the LDS crossbar (ds_bpermute): DPP quad permutations pair lanes i^1, i^2;
// row half-mirror / mirror pair lane i with the other quad of its 8 / the
// other 8 of its 16-lane row; v_permlane16_swap / v_permlane32_swap exchange
// odd and even 16-lane rows / the two 32-lane halves. Every step combines
// two lanes of the group, both lanes compute a + b with the same operands
// (addition commutes), so the whole group ends with the bitwise same value.
template <int CTRL, int BANK = 0xF>
__device__ __forceinline__ float dpp_f(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xF, BANK, false));
}
// Every lane written, from a source lane inside its row for each control
// used with it (quad permutations, mirrors, row broadcast), so no old value
// is needed: the compiler folds the move into the consuming add
// (v_add_f32_dpp), one instruction per reduction level instead of three.
template <int CTRL>
__device__ __forceinline__ float dpp_all(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int2 w = *reinterpret_cast<const int2*>(&v);
  int2 r;
  r.x = __builtin_amdgcn_mov_dpp(w.x, CTRL, 0xF, 0xF, true);
  r.y = __builtin_amdgcn_mov_dpp(w.y, CTRL, 0xF, 0xF, true);
  return *reinterpret_cast<const double*>(&r);
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7-i within 8
constexpr int DPP_MIRROR = 0x140;      // lane i <-> 15-i within 16
constexpr int DPP_QUAD_BCAST0 = 0x00;  // quad_perm [0,0,0,0]
constexpr int DPP_PAIR_BCAST0 = 0xA0;  // quad_perm [0,0,2,2]
constexpr int DPP_ROW_SHR4 = 0x114;
constexpr int DPP_ROW_BCAST0 = 0x150;  // row_newbcast:0

// (a, b) after swapping the odd 16-lane rows of a with the even rows of b,
// called with a = b = v: first = rows (0,0,2,2), second = rows (1,1,3,3).
__device__ __forceinline__ void swap16(float v, float& lo, float& hi) {
  auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  lo = __uint_as_float(r[0]);
  hi = __uint_as_float(r[1]);
}
// The same over 32-lane halves: first = lanes 0-31 twice, second = 32-63 twice.
__device__ __forceinline__ void swap32(float v, float& lo, float& hi) {
  auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  lo = __uint_as_float(r[0]);
  hi = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap16(double v, double& lo, double& hi) {
  float a, b, c, e;
  const float2 w = *reinterpret_cast<const float2*>(&v);
  swap16(w.x, a, b);
  swap16(w.y, c, e);
  float2 l{a, c}, h{b, e};
  lo = *reinterpret_cast<const double*>(&l);
  hi = *reinterpret_cast<const double*>(&h);
}
__device__ __forceinline__ void swap32(double v, double& lo, double& hi) {
  float a, b, c, e;
  const float2 w = *reinterpret_cast<const float2*>(&v);
  swap32(w.x, a, b);
  swap32(w.y, c, e);
  float2 l{a, c}, h{b, e};
  lo = *reinterpret_cast<const double*>(&l);
  hi = *reinterpret_cast<const double*>(&h);
}

__device__ __forceinline__ float group_sum(float v, int lpc) {
  if (lpc >= 2) v += dpp_all<DPP_XOR1>(v);
  if (lpc >= 4) v += dpp_all<DPP_XOR2>(v);
  if (lpc >= 8) v += dpp_all<DPP_HALF_MIRROR>(v);
  if (lpc >= 16) v += dpp_all<DPP_MIRROR>(v);
  float lo, hi;
  if (lpc >= 32) { swap16(v, lo, hi); v = lo + hi; }
  if (lpc >= 64) { swap32(v, lo, hi); v = lo + hi; }
  return v;
}
__device__ __forceinline__ double group_sum(double v, int lpc) {
  if (lpc >= 2) v += dpp_d<DPP_XOR1>(v);
  if (lpc >= 4) v += dpp_d<DPP_XOR2>(v);
  if (lpc >= 8) v += dpp_d<DPP_HALF_MIRROR>(v);
  if (lpc >= 16) v += dpp_d<DPP_MIRROR>(v);
  double lo, hi;
  if (lpc >= 32) { swap16(v, lo, hi); v = lo + hi; }
  if (lpc >= 64) { swap32(v, lo, hi); v = lo + hi; }
  return v;
}
__device__ __forceinline__ float group_max(float v, int lpc) {
  if (lpc >= 2) v = fmaxf(v, dpp_all<DPP_XOR1>(v));
  if (lpc >= 4) v = fmaxf(v, dpp_all<DPP_XOR2>(v));
  if (lpc >= 8) v = fmaxf(v, dpp_all<DPP_HALF_MIRROR>(v));
  if (lpc >= 16) v = fmaxf(v, dpp_all<DPP_MIRROR>(v));
  float lo, hi;
  if (lpc >= 32) { swap16(v, lo, hi); v = fmaxf(lo, hi); }
  if (lpc >= 64) { swap32(v, lo, hi); v = fmaxf(lo, hi); }
  return v;
}
// The value of the group's first lane, in every lane of the group.
__device__ __forceinline__ float group_bcast0(float v, int lpc) {
  if (lpc == 1) return v;
  if (lpc == 2) return dpp_all<DPP_PAIR_BCAST0>(v);
  if (lpc == 4) return dpp_all<DPP_QUAD_BCAST0>(v);
  if (lpc == 8) {
    const float q = dpp_all<DPP_QUAD_BCAST0>(v);
    return dpp_f<DPP_ROW_SHR4, 0xA>(q, q);  // banks 1 and 3 take lane i-4
  }
  if (lpc == 64) return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));  // int builtin: bitcast
  const float rb = dpp_all<DPP_ROW_BCAST0>(v);
  if (lpc == 16) return rb;
  float lo, hi;
  swap16(rb, lo, hi);
  return lo;
}

__device__ __forceinline__ float& f4(float4& v, int e) { return (&v.x)[e]; }
__device__ __forceinline__ float f4c(const float4& v, int e) { return (&v.x)[e]; }

}  // namespace fmc
