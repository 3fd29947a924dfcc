// Wave-level helpers shared by the libabea kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abea {

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Ordering of this wave's own LDS and global accesses across its lanes (all
// lanes of a wave belong to one workgroup and share the CU's caches).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Ordering of this wave's LDS accesses across its lanes.  The LDS executes a
// wave's operations in issue order, so only the compiler needs the fence.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Number of lanes below this one whose bit is set in mask.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t base_code(char c) {
  // GPU/align_cuda.cl:22-35: A,C,G,T -> 0..3, anything else -> 0
  return c == 'C' ? 1u : c == 'G' ? 2u : c == 'T' ? 3u : 0u;
}

// Wave-wide scans and reductions on DPP (row shifts within 16-lane rows, then
// the row broadcasts of lanes 15 and 31), no LDS round trips.

// inclusive prefix sum over the 64 lanes
__device__ __forceinline__ int wave_scan_add(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

__device__ __forceinline__ double dpp_f64(double v, int ctrl, int row_mask) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  // the row broadcasts need a per-call immediate, so spell the cases out
  uint32_t lo, hi;
#define ABEA_DPP64(C_, R_)                                                            \
  lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, C_, R_, 0xf, false); \
  hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), C_, R_, 0xf, false)
  switch (ctrl) {
    case 0x111: ABEA_DPP64(0x111, 0xf); break;
    case 0x112: ABEA_DPP64(0x112, 0xf); break;
    case 0x114: ABEA_DPP64(0x114, 0xf); break;
    case 0x118: ABEA_DPP64(0x118, 0xf); break;
    case 0x142: ABEA_DPP64(0x142, 0xa); break;
    default: ABEA_DPP64(0x143, 0xc); break;
  }
#undef ABEA_DPP64
  (void)row_mask;
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// sum over the 64 lanes (lane 63 of an inclusive scan), broadcast.  Only for
// sums whose every partial sum is exact: the order of additions differs
// from a serial sum.
__device__ __forceinline__ double wave_sum_exact_f64(double v) {
  v += dpp_f64(v, 0x111, 0xf);
  v += dpp_f64(v, 0x112, 0xf);
  v += dpp_f64(v, 0x114, 0xf);
  v += dpp_f64(v, 0x118, 0xf);
  v += dpp_f64(v, 0x142, 0xa);
  v += dpp_f64(v, 0x143, 0xc);
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// maximum over the 64 lanes of non-negative floats, broadcast
__device__ __forceinline__ float wave_max_nonneg(float v) {
  int x = __builtin_bit_cast(int, v);  // non-negative floats order as integers
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(x, 63));
}

// lane j's value of a double, broadcast
__device__ __forceinline__ double readlane_f64(double v, int j) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), j);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// inclusive prefix maximum over the 64 lanes (identity: INT32_MIN)
__device__ __forceinline__ int wave_scan_max(int v) {
  const int lo = (int)0x80000000;
  v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

}  // namespace abea
