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

// Bases k..k+5 of the read (k in [0, len - 6]): a 4-byte and a 2-byte load
// at the bases' own (unaligned) address, both inside the read.
__device__ __forceinline__ void kmer_bytes(const char *S, int k, uint32_t &lo, uint32_t &hi) {
  uint16_t h;
  __builtin_memcpy(&lo, S + k, 4);
  __builtin_memcpy(&h, S + k + 4, 2);
  hi = h;
}

// Base codes of four bases packed in a word (A,C,G,T -> 0..3, anything else
// -> 0, as GPU/align_cuda.cl:22-35), two bits per byte, without branches:
// ((c >> 1) ^ (c >> 2)) & 3 is the code of A, C, G and T, and an exact
// per-byte zero test of c ^ base keeps it only for those four.
__device__ __forceinline__ uint32_t zero_bytes(uint32_t y) {
  return ~(((y & 0x7f7f7f7fu) + 0x7f7f7f7fu) | y | 0x7f7f7f7fu);  // 0x80 where a byte is 0
}

__device__ __forceinline__ uint32_t base_codes4(uint32_t x) {
  const uint32_t code = ((x >> 1) ^ (x >> 2)) & 0x03030303u;
  const uint32_t ok = zero_bytes(x ^ 0x41414141u) | zero_bytes(x ^ 0x43434343u) |
                      zero_bytes(x ^ 0x47474747u) | zero_bytes(x ^ 0x54545454u);
  return code & ((ok >> 7) * 3u);
}

// 6-mer rank of bases 0-3 (lo) and 4-5 (hi), first base most significant
// (GPU/align_cuda.cl:39-51); the multiply gathers the four 2-bit codes of lo
// into bits 24-31 in reverse byte order.
__device__ __forceinline__ uint32_t rank_lohi(uint32_t lo, uint32_t hi) {
  const uint32_t c4 = base_codes4(lo), c2 = base_codes4(hi);
  return (((c4 * 0x40100401u) >> 24) << 4) | ((c2 & 3u) << 2) | ((c2 >> 8) & 3u);
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
