// Device-side helpers for gfx950 (CDNA4, wave64): lane masks, wave prefix sums and
// device-scope atomics used by the look-back kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gprt {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// The lanes of `lanes` whose BITS-bit digit equals this lane's; every lane of `lanes` must be
// active. One ballot per digit bit; a lane differs from this one where the ballot's bit differs
// from this lane's bit s (sign-extended to 0 or ~0), so the complement of the result collects
// ballot ^ s over the bits: one v_bitop3_b32 per 32-lane half and bit (truth table 0xF6 =
// S0 | (S1 ^ S2), index S0 * 4 + S1 * 2 + S2), after a v_bfe_i32 and the ballot's compare. About
// 4 vector instructions per bit, where selecting ballot or ~ballot per lane as 64-bit values
// compiled to about 10.
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t lanes) {
  uint32_t lo = ~(uint32_t)lanes, hi = ~(uint32_t)(lanes >> 32);
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint32_t s = (uint32_t)__builtin_amdgcn_sbfe((int)d, (unsigned)b, 1u);
    const uint64_t bb = __builtin_amdgcn_uicmp(s, 0u, 33);  // 33: not equal; inactive lanes read 0
    lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)bb, s, 0xF6);
    hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(bb >> 32), s, 0xF6);
  }
  return ~(((uint64_t)hi << 32) | lo);
}

// Inclusive prefix sum across the 64 lanes with DPP: shifts of 1, 2, 4 and 8 lanes within each
// 16-lane row (lanes shifted in from outside the row read 0), then lane 15 of each row is added
// to the next row and lane 31 to the upper half. Six VALU operations with no LDS traffic; the
// whole wave must be active.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return v;
}

__device__ __forceinline__ uint32_t wave_read(uint32_t v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Device-scope (agent) relaxed atomics: the value carried by a look-back word is the payload,
// so no additional ordering is required between the word and other memory.
__device__ __forceinline__ uint64_t ld_agent(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ticket(uint32_t *p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Order-preserving float <-> uint mapping (for exact, order-free min/max via integer atomics).
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

}  // namespace gprt
