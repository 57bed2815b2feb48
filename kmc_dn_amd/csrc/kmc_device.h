// Device-side building blocks shared by the KMC kernels (gfx950, wave64).
//
//  * Philox4x32-10 counter-based RNG and the engine's stream contract:
//    key = trajectory seed, counter = (hop lo, hop hi, 0, 0); words 0,1 give
//    the clock uniform, words 2,3 the event uniform (53-bit, [0,1)).
//  * Wave-wide inclusive scans in FP64 and FP32 built on DPP row shifts and
//    row broadcasts (no LDS round trip).
//  * Ballot / mbcnt / readlane helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmc_types.h"

namespace kmc {

// ----------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ----------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // each 32 x 32 -> 64-bit product in one multiply (v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Uniforms of hop `hop` of the stream keyed by `seed`.
__device__ __forceinline__ void stream_uniforms(uint64_t seed, uint64_t hop, double& u_clock,
                                                double& u_event) {
  const u32x4 o = philox4x32_10(u32x4{(uint32_t)hop, (uint32_t)(hop >> 32), 0u, 0u},
                                (uint32_t)seed, (uint32_t)(seed >> 32));
  u_clock = u53(o.x, o.y);
  u_event = u53(o.z, o.w);
}

// ----------------------------------------------------------------------------
// Wave helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ int mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int readlane_i32(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const uint64_t b = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Orders this wave's LDS accesses (writes by some lanes, reads by others).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP controls (GFX9 encoding).
enum : int {
  DPP_ROW_SHR1 = 0x111,
  DPP_ROW_SHR2 = 0x112,
  DPP_ROW_SHR4 = 0x114,
  DPP_ROW_SHR8 = 0x118,
  DPP_ROW_BCAST15 = 0x142,
  DPP_ROW_BCAST31 = 0x143,
};

template <int CTRL, int ROW_MASK, bool BOUND>
__device__ __forceinline__ float dpp_f32(float x) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xF, BOUND));
}

template <int CTRL, int ROW_MASK, bool BOUND>
__device__ __forceinline__ double dpp_f64(double x) {
  const uint64_t b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, BOUND);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xF, BOUND);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// DPP move that writes only the rows of ROW_MASK; the other rows keep `old`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64_rows(double old, double x) {
  const uint64_t b = __double_as_longlong(x), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(o >> 32), (int)(uint32_t)(b >> 32), CTRL,
                                             ROW_MASK, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Inclusive scan over the 64 lanes of a wave (Hillis-Steele inside each row of
// 16 lanes, then row broadcasts).  Every lane must be active.
__device__ __forceinline__ double wave_scan_f64(double x) {
  x += dpp_f64<DPP_ROW_SHR1, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_SHR2, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_SHR4, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_SHR8, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_BCAST15, 0xA, false>(x);
  x += dpp_f64<DPP_ROW_BCAST31, 0xC, false>(x);
  return x;
}

// The same scan for loops: the row broadcasts land in two carried registers
// whose rows outside each broadcast's row mask are zero (rows 0 and 2 of b15,
// rows 0 and 1 of b31) and are never written, so no per-call zeroing is needed.
// Both start as 0.0 and are passed to every call.
__device__ __forceinline__ double wave_scan_f64(double x, double& b15, double& b31) {
  x += dpp_f64<DPP_ROW_SHR1, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_SHR2, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_SHR4, 0xF, true>(x);
  x += dpp_f64<DPP_ROW_SHR8, 0xF, true>(x);
  b15 = dpp_f64_rows<DPP_ROW_BCAST15, 0xA>(b15, x);
  x += b15;
  b31 = dpp_f64_rows<DPP_ROW_BCAST31, 0xC>(b31, x);
  x += b31;
  return x;
}

// The row broadcasts add in place with one DPP add each: rows outside the row
// mask are not written and keep x, which is what x + 0 gives (the compiler's
// move-then-add form needs a zeroed register and two more instructions).  The
// s_nop gives the DPP read of x the two wait states it needs after the VALU write.
__device__ __forceinline__ float wave_scan_f32(float x) {
  x += dpp_f32<DPP_ROW_SHR1, 0xF, true>(x);
  x += dpp_f32<DPP_ROW_SHR2, 0xF, true>(x);
  x += dpp_f32<DPP_ROW_SHR4, 0xF, true>(x);
  x += dpp_f32<DPP_ROW_SHR8, 0xF, true>(x);
  asm volatile(
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(x));
  return x;
}

// Inclusive max-scan of unsigned 32-bit values over the 64 lanes (the DPP row
// shifts and row broadcasts of wave_scan_f64, with 0 as the identity).  Every
// lane must be active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
  auto step = [](uint32_t v, uint32_t s) { return v > s ? v : s; };
  x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR1, 0xF, 0xF, true));
  x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR2, 0xF, 0xF, true));
  x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR4, 0xF, 0xF, true));
  x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR8, 0xF, 0xF, true));
  x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST15, 0xA, 0xF, false));
  x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST31, 0xC, 0xF, false));
  return x;
}

// Maximum over the 64 lanes (every lane active), the same on every lane: the
// row shifts and row broadcasts of the scans with max; a lane with no source
// lane keeps its own value (bound_ctrl off, `old` = x), which max leaves alone.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_keep_f64(double x) {
  const uint64_t b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)b, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(b >> 32), (int)(uint32_t)(b >> 32), CTRL,
                                             ROW_MASK, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
__device__ __forceinline__ double wave_max_f64(double x) {
  x = fmax(x, dpp_keep_f64<DPP_ROW_SHR1, 0xF>(x));
  x = fmax(x, dpp_keep_f64<DPP_ROW_SHR2, 0xF>(x));
  x = fmax(x, dpp_keep_f64<DPP_ROW_SHR4, 0xF>(x));
  x = fmax(x, dpp_keep_f64<DPP_ROW_SHR8, 0xF>(x));
  x = fmax(x, dpp_keep_f64<DPP_ROW_BCAST15, 0xA>(x));
  x = fmax(x, dpp_keep_f64<DPP_ROW_BCAST31, 0xC>(x));
  return readlane_f64(x, 63);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_keep_f32(float x) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, ROW_MASK, 0xF, false));
}
__device__ __forceinline__ float wave_max_f32(float x) {
  x = fmaxf(x, dpp_keep_f32<DPP_ROW_SHR1, 0xF>(x));
  x = fmaxf(x, dpp_keep_f32<DPP_ROW_SHR2, 0xF>(x));
  x = fmaxf(x, dpp_keep_f32<DPP_ROW_SHR4, 0xF>(x));
  x = fmaxf(x, dpp_keep_f32<DPP_ROW_SHR8, 0xF>(x));
  x = fmaxf(x, dpp_keep_f32<DPP_ROW_BCAST15, 0xA>(x));
  x = fmaxf(x, dpp_keep_f32<DPP_ROW_BCAST31, 0xC>(x));
  return readlane_f32(x, 63);
}

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }

// Occupation bitmask of up to 256 acceptors as four named words.  Lane-varying
// bit tests select among SSA values; a private uint64_t[4] indexed that way is
// turned into scratch-memory loads by the compiler.
struct Occ4 {
  uint64_t a, b, c, d;
  __device__ __forceinline__ uint64_t word(int w) const {
    return w == 0 ? a : (w == 1 ? b : (w == 2 ? c : d));
  }
  __device__ __forceinline__ bool test(int i) const { return (word(i >> 6) >> (i & 63)) & 1ull; }
  __device__ __forceinline__ void assign(int i, bool v) {
    const uint64_t bit = 1ull << (i & 63);
    const int w = i >> 6;
    if (w == 0) a = v ? (a | bit) : (a & ~bit);
    else if (w == 1) b = v ? (b | bit) : (b & ~bit);
    else if (w == 2) c = v ? (c | bit) : (c & ~bit);
    else d = v ? (d | bit) : (d & ~bit);
  }
};

}  // namespace kmc
