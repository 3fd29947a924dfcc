// ABEA kernels for gfx950 (CDNA4, wave64).
//
//   pre      stage per-k-mer emission parameters and SoA event means
//            (replaces align_kernel_pre_2d, GPU/align_cuda.cl:162-276)
//   core     adaptive band walk, one wavefront per read
//            (replaces align_kernel_core_2d_shm, GPU/align_cuda.cl:297-609)
//   post     best end cell, traceback over the uint8 trace, QC
//            (replaces align_kernel_post, GPU/align_cuda.cl:618-899)
//   adjoint  gradient of the best-path score along the traced path (new)
//
// Band representation in the core.  Band b holds cells (e, k) with
// e + k = b - 2 and k in [K_b, K_b + 99] (K_b = lower-left k-mer of the band,
// GPU/align_cuda.cl:112-117).  A wave keeps the band in 128 "slots", two per
// lane, and the slot of a cell is k mod 128.  With that indexing the three
// DP neighbours never depend on the band moves:
//   up   (e-1, k)   = band b-1, same slot
//   left (e,   k-1) = band b-1, slot - 1
//   diag (e-1, k-1) = band b-2, slot - 1 = the "left" vector of band b-1
// so one DPP rotate per band supplies them, each lane's k-mer parameters stay
// put (a right move refills one slot), and the event under every slot shifts
// by one slot per band (a down move feeds one new event in).  Slots outside
// the band window hold -inf, so out-of-band neighbours read -inf exactly as
// the reference's is_offset_valid() tests do.
//
// Numerics follow f5c's CPU align(): float cells, emission in float with an
// IEEE divide and no contraction (built with -ffp-contract=off), transition
// sums in double (the reference's constants are double, GPU/align_cuda.cl:
// 355-359), argmax with the reference's tie order (GPU/align_cuda.cl:536-542).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "abea_internal.h"

namespace abea {

#define ABEA_NEG_INF (-__builtin_inff())

typedef __attribute__((address_space(4))) const float c_f32;
typedef __attribute__((address_space(4))) const int32_t c_i32;
typedef __attribute__((address_space(4))) const uint32_t c_u32;

constexpr uint32_t FROM_D = 0, FROM_U = 1, FROM_L = 2;  // GPU/align_cuda.cl:130-132

// Uniform (scalar-cache) load of one read's metadata record.
__device__ __forceinline__ ReadMeta load_meta(const ReadMeta *meta, int r) {
  const c_i32 *p = (const c_i32 *)(meta + r);
  int32_t w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = p[i];
  ReadMeta m;
  __builtin_memcpy(&m, w, sizeof(m));
  return m;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// lane l receives lane (l - 1) mod 64: DPP wave_ror:1.
__device__ __forceinline__ float rotate_up(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x13C, 0xF, 0xF, false));
}

// Value of slot j (0..127) of a band held as (slot 2l, slot 2l+1) per lane.
__device__ __forceinline__ float slot_value(float s0, float s1, int j) {
  const int a = __builtin_amdgcn_readlane(__builtin_bit_cast(int, s0), j >> 1);
  const int b = __builtin_amdgcn_readlane(__builtin_bit_cast(int, s1), j >> 1);
  return __builtin_bit_cast(float, (j & 1) ? b : a);
}

__device__ __forceinline__ uint32_t base_code(char c) {
  // GPU/align_cuda.cl:22-35: A,C,G,T -> 0..3, anything else -> 0
  return c == 'C' ? 1u : c == 'G' ? 2u : c == 'T' ? 3u : 0u;
}

// f5c emission, GPU/align_cuda.cl:511-528: a = (x - m) / s; em = c + (-0.5a)a
// with c = -0.918938f - level_log_stdv precomputed by pre.
__device__ __forceinline__ float emission(float x, float m, float s, float c) {
  const float a = (x - m) / s;
  return c + ((-0.5f * a) * a);
}

// One DP cell, GPU/align_cuda.cl:532-542 with double transition constants.
__device__ __forceinline__ void dp_cell(float diag, float up, float left, float em,
                                        double lp_step, double lp_stay, double lp_skip,
                                        float &score, uint32_t &from) {
  const double emd = (double)em;
  const float sd = (float)(((double)diag + lp_step) + emd);
  const float su = (float)(((double)up + lp_stay) + emd);
  const float sl = (float)((double)left + lp_skip);
  float mx = sd;
  uint32_t f = FROM_D;
  mx = su > mx ? su : mx;
  f = mx == su ? FROM_U : f;
  mx = sl > mx ? sl : mx;
  f = mx == sl ? FROM_L : f;
  score = mx;
  from = f;
}

// ---------------------------------------------------------------------------
// pre: one workgroup per read (grid-stride), coalesced over k-mers and events.
__global__ __launch_bounds__(256) void abea_pre_kernel(KernelArgs a, int q_begin, int q_count) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0;  // persistent core's queue head
  for (int qi = blockIdx.x; qi < q_count; qi += gridDim.x) {
    const int r = ((c_i32 *)a.queue)[q_begin + qi];
    const ReadMeta m = load_meta(a.meta, r);
    if (m.n_k <= 0 || m.n_e <= 0) continue;
    const char *s = a.seq + a.seq_ptr[r];
    const abea_scalings_t sc = a.scalings[r];
    KmerParam *km = a.km + m.km_off;
    for (int k = threadIdx.x; k < m.n_k; k += blockDim.x) {
      uint32_t rank = 0;
#pragma unroll
      for (int i = 0; i < ABEA_KMER_SIZE; ++i) rank = (rank << 2) | base_code(s[k + i]);
      const abea_model_t md = a.model[rank];
      KmerParam p;
      p.mean = sc.scale * md.level_mean + sc.shift;  // two roundings, as the reference
      p.stdv = md.level_stdv;                        // scaling.var unused, GPU/align_cuda.cl:515
      p.c = -0.918938f - md.level_log_stdv;          // cached log, GPU/align_cuda.cl:517-518
      p.rank = (int32_t)rank;
      km[k] = p;
    }
    float *ev = a.ev + m.ev_off;
    float *endcol = a.endcol + m.ev_off;
    for (int e = threadIdx.x; e < m.n_e; e += blockDim.x) {
      if (a.events) ev[e] = a.events[m.ev_off + e].mean;
      endcol[e] = ABEA_NEG_INF;
    }
  }
}

// ---------------------------------------------------------------------------
// core: the band walk of one read by one wavefront.
__device__ __forceinline__ void core_read(const KernelArgs &a, int r, int lane) {
  const ReadMeta m = load_meta(a.meta, r);
  const int n_e = m.n_e, n_k = m.n_k;
  if (n_k <= 0 || n_e <= 0) return;
  const int n_bands = n_e + n_k + 2;
  const double lp_stay = m.lp_stay, lp_step = m.lp_step;
  const double lp_skip = a.lp_skip, lp_trim = a.lp_trim;
  const float *evp = a.ev + m.ev_off;
  const KmerParam *kmp = a.km + m.km_off;
  c_f32 *evs = (c_f32 *)evp;
  c_f32 *kms = (c_f32 *)kmp;
  float *endcol = a.endcol + m.ev_off;
  uint8_t *trace = a.trace + m.trace_off;
  uint32_t *moves = a.moves + m.mv_off;
  c_u32 *replay = a.replay ? (c_u32 *)(a.replay + m.mv_off) : nullptr;

  const int j0 = 2 * lane, j1 = 2 * lane + 1;
  // band 1 lower-left (E, K) = (50, -51), GPU/align_cuda.cl:251-255
  int E = ABEA_BANDWIDTH / 2, K = -1 - ABEA_BANDWIDTH / 2;

  // Event and k-mer parameters under each slot for band 1's window.
  float ev0, ev1, km0m, km0s, km0c, km1m, km1s, km1c;
  {
    const int o0 = (j0 - K) & 127, o1 = (j1 - K) & 127;
    const int e0 = E - o0, e1 = E - o1;
    ev0 = (e0 >= 0 && e0 < n_e) ? evp[e0] : 0.f;
    ev1 = (e1 >= 0 && e1 < n_e) ? evp[e1] : 0.f;
    const int k0 = K + o0, k1 = K + o1;
    KmerParam p0{0.f, 1.f, 0.f, 0}, p1{0.f, 1.f, 0.f, 0};
    if (k0 >= 0 && k0 < n_k) p0 = kmp[k0];
    if (k1 >= 0 && k1 < n_k) p1 = kmp[k1];
    km0m = p0.mean; km0s = p0.stdv; km0c = p0.c;
    km1m = p1.mean; km1s = p1.stdv; km1c = p1.c;
  }
  // band 1: only the first trim cell (e=0, k=-1 -> slot 127) = lp_trim
  // (GPU/align_cuda.cl:263-267); band 0: start cell (e=-1,k=-1) = 0, whose
  // left-shifted copy is the diag source of band 2.
  float b1_0 = ABEA_NEG_INF, b1_1 = (lane == 63) ? (float)lp_trim : ABEA_NEG_INF;
  float lv0 = (lane == 0) ? 0.0f : ABEA_NEG_INF, lv1 = ABEA_NEG_INF;

  // scalar prefetch of what band 2 may need: next event, k-mer K+128
  float nx_ev = evs[min(E + 1, n_e - 1)];
  int kn = min(K + 128, n_k - 1);
  float nk_m = kms[4 * kn], nk_s = kms[4 * kn + 1], nk_c = kms[4 * kn + 2];

  const int js = (n_k - 1) & 127;  // slot of the last k-mer (end-cell column)
  uint32_t mv = 0;
  uint16_t *row = (uint16_t *)(trace + 2 * kRowBytes) + lane;

  for (int b = 2; b < n_bands; ++b, row += kRowBytes / 2) {
    // Suzuki's rule on band b-1's lower-left and upper-right cells,
    // GPU/align_cuda.cl:398-410.
    const float ll = slot_value(b1_0, b1_1, K & 127);
    const float ur = slot_value(b1_0, b1_1, (K + ABEA_BANDWIDTH - 1) & 127);
    bool right;
    if (replay)
      right = (replay[b >> 5] >> (b & 31)) & 1u;
    else if (ll == ABEA_NEG_INF && ur == ABEA_NEG_INF)
      right = (b & 1) != 0;
    else
      right = ll < ur;
    mv |= (uint32_t)right << (b & 31);
    if ((b & 31) == 31) {
      if (lane == 0) moves[b >> 5] = mv;
      mv = 0;
    }

    const float e_rot = rotate_up(ev1);
    const float l_rot = rotate_up(b1_1);
    float nev0 = e_rot, nev1 = ev0;
    if (right) {
      K += 1;
      const int jn = (K + 127) & 127;  // slot that now holds k-mer K+127
      const bool mine = lane == (jn >> 1);
      if (jn & 1) {
        km1m = mine ? nk_m : km1m;
        km1s = mine ? nk_s : km1s;
        km1c = mine ? nk_c : km1c;
      } else {
        km0m = mine ? nk_m : km0m;
        km0s = mine ? nk_s : km0s;
        km0c = mine ? nk_c : km0c;
      }
    } else {
      E += 1;
      const int jb = K & 127;  // bottom slot takes the new event E
      const bool mine = lane == (jb >> 1);
      if (jb & 1)
        nev1 = mine ? nx_ev : nev1;
      else
        nev0 = mine ? nx_ev : nev0;
    }
    ev0 = nev0;
    ev1 = nev1;
    nx_ev = evs[min(E + 1, n_e - 1)];
    kn = min(K + 128, n_k - 1);
    nk_m = kms[4 * kn];
    nk_s = kms[4 * kn + 1];
    nk_c = kms[4 * kn + 2];

    // fill range of offsets, GPU/align_cuda.cl:460-469
    const int lo = max(max(-K, E - (n_e - 1)), 0);
    const int hi = min(min(n_k - K, E + 1), ABEA_BANDWIDTH);
    const uint32_t span = hi > lo ? (uint32_t)(hi - lo) : 0u;
    const int o0 = (j0 - K) & 127, o1 = (j1 - K) & 127;
    const bool v0 = (uint32_t)(o0 - lo) < span;
    const bool v1 = (uint32_t)(o1 - lo) < span;

    float M0, M1;
    uint32_t f0, f1;
    dp_cell(lv0, b1_0, l_rot, emission(ev0, km0m, km0s, km0c), lp_step, lp_stay, lp_skip, M0, f0);
    dp_cell(lv1, b1_1, b1_0, emission(ev1, km1m, km1s, km1c), lp_step, lp_stay, lp_skip, M1, f1);
    M0 = v0 ? M0 : ABEA_NEG_INF;
    M1 = v1 ? M1 : ABEA_NEG_INF;
    f0 = v0 ? f0 : 0u;
    f1 = v1 ? f1 : 0u;
    if (K < 0) {
      // trim cell (e, k=-1) sits in slot 127 while K <= -1,
      // GPU/align_cuda.cl:432-443
      const int et = E - (-1 - K);
      const bool in = et >= 0 && et < n_e;
      const float tv = in ? (float)(lp_trim * (double)(et + 1)) : ABEA_NEG_INF;
      if (lane == 63) {
        M1 = tv;
        f1 = in ? FROM_U : 0u;
      }
    }
    *row = (uint16_t)(f0 | (f1 << 8));

    // end-cell column (k = n_k - 1) for the post kernel's end search
    const int os = (n_k - 1) - K;  // offset of the last k-mer in this band
    if ((uint32_t)(os - lo) < span) {
      const float v = (js & 1) ? M1 : M0;
      if (lane == (js >> 1)) endcol[E - os] = v;
    }

    lv0 = l_rot;
    lv1 = b1_0;
    b1_0 = M0;
    b1_1 = M1;
  }
  if (((n_bands - 1) & 31) != 31 && lane == 0) moves[(n_bands - 1) >> 5] = mv;
}

template <int WPB>
__global__ __launch_bounds__(64 * WPB) void abea_core_kernel(KernelArgs a, int q_begin, int q_count,
                                                             int persistent) {
  const int lane = threadIdx.x & 63;
  int qi;
  if (persistent) {
    int t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1);
    qi = uniform(t);
  } else {
    qi = uniform(blockIdx.x * WPB + (threadIdx.x >> 6));
  }
  while (qi < q_count) {
    const int r = ((c_i32 *)a.queue)[q_begin + qi];
    core_read(a, r, lane);
    if (!persistent) break;
    int t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1);
    qi = uniform(t);
  }
}

// ---------------------------------------------------------------------------
// post: one wavefront per read.
//  1. best end cell on the last k-mer (GPU/align_cuda.cl:722-746), wave-wide
//     with the reference's first-maximum rule
//  2. traceback (GPU/align_cuda.cl:755-828): the trace rows the path visits
//     are staged through LDS in chunks of ROWS rows; lane 0 follows the
//     from-codes and records them in traceback order
//  3. pairs in ascending order (no reversal pass): the path is rebuilt from
//     its end cell with a wave prefix sum over the recorded moves; the QC
//     emission sum runs in the reference's (traceback) order, in double
//  4. QC (GPU/align_cuda.cl:856-882)
template <int ROWS>
__global__ __launch_bounds__(64) void abea_post_kernel(KernelArgs a, int q_begin, int q_count) {
  __shared__ uint4 tile[ROWS * (kRowBytes / 16)];
  __shared__ float em_buf[64];
  const int lane = threadIdx.x;
  const uint8_t *tile_b = (const uint8_t *)tile;
  for (int qi = blockIdx.x; qi < q_count; qi += gridDim.x) {
    const int r = ((c_i32 *)a.queue)[q_begin + qi];
    const ReadMeta m = load_meta(a.meta, r);
    const int n_e = m.n_e, n_k = m.n_k;
    const abea_output_t &o = a.out;
    float best = ABEA_NEG_INF;
    int best_e = 0x7fffffff;
    if (n_k > 0 && n_e > 0) {
      const float *endcol = a.endcol + m.ev_off;
      for (int e = lane; e < n_e; e += 64) {
        const float s = (float)((double)endcol[e] + (double)(n_e - e) * a.lp_trim);
        if (s > best) {
          best = s;
          best_e = e;
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const int oe = __shfl_xor(best_e, off);
        if (ob > best || (ob == best && oe < best_e)) {
          best = ob;
          best_e = oe;
        }
      }
    }
    if (n_k <= 0 || n_e <= 0 || best == ABEA_NEG_INF) {
      // empty read, or the band never reached the last k-mer (the reference
      // would trace cells outside the band): reported as a failed read
      if (lane == 0) {
        o.n_pairs[r] = 0;
        if (o.path_len) o.path_len[r] = 0;
        if (o.score) o.score[r] = best;
        if (o.avg_log_emission) o.avg_log_emission[r] = 0.0;
      }
      continue;
    }

    // ---- traceback
    uint8_t *rev = a.rev_from + m.pair_off;
    const uint8_t *trace = a.trace + m.trace_off;
    int k = n_k - 1, e = uniform(best_e);
    int b = e + k + 2;
    int n = 0, gap = 0, max_gap = 0;
    int lo = b - ROWS + 1;
    for (;;) {
      const int r0 = max(lo, 0);
      const uint4 *src = (const uint4 *)(trace + (int64_t)r0 * kRowBytes);
      const int nvec = (b - r0 + 1) * (kRowBytes / 16);
      const int base = (r0 - lo) * (kRowBytes / 16);
      for (int i = lane; i < nvec; i += 64) tile[base + i] = src[i];
      __syncthreads();
      if (lane == 0) {
        while (b >= lo) {
          const uint32_t f = tile_b[(b - lo) * kRowBytes + (k & 127)];
          rev[n++] = (uint8_t)f;
          if (f == FROM_D) {
            --k;
            --e;
            b -= 2;
            gap = 0;
          } else if (f == FROM_U) {
            --e;
            --b;
            gap = 0;
          } else {
            --k;
            --b;
            ++gap;
            max_gap = max(max_gap, gap);
          }
          if (k < 0 || e < 0) break;
        }
      }
      __syncthreads();
      k = __shfl(k, 0);
      e = __shfl(e, 0);
      b = __shfl(b, 0);
      n = __shfl(n, 0);
      if (k < 0 || e < 0) break;
      lo -= ROWS;
    }
    max_gap = __shfl(max_gap, 0);
    const int P = n;
    __threadfence();  // lane 0's from-codes become visible to the whole wave

    // ---- ascending pairs, from the end cell down, 64 pairs per step
    const float *evp = a.ev + m.ev_off;
    const KmerParam *kmp = a.km + m.km_off;
    abea_aligned_pair_t *pairs = o.pairs + m.pair_off;
    uint8_t *from_out = o.path_from ? o.path_from + m.pair_off : nullptr;
    int ck = n_k - 1, ce = uniform(best_e);  // position of the chunk's top pair
    double sum_em = 0.0;
    int first_k = 0;
    for (int t0 = 0; t0 < P; t0 += 64) {
      const int t = t0 + lane;  // t-th pair in traceback order = ascending index P-1-t
      const bool act = t < P;
      const uint32_t f = act ? rev[t] : FROM_D;
      // move from ascending pair i-1 to pair i (= inverse of the from code)
      const int dk = act && f != FROM_U ? 1 : 0;
      const int de = act && f != FROM_L ? 1 : 0;
      int pk = dk, pe = de;  // inclusive prefix over lanes
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int xk = __shfl_up(pk, off);
        const int xe = __shfl_up(pe, off);
        if (lane >= off) {
          pk += xk;
          pe += xe;
        }
      }
      const int kk = ck - (pk - dk);
      const int ee = ce - (pe - de);
      if (act) {
        const int i = P - 1 - t;
        pairs[i].ref_pos = kk;
        pairs[i].read_pos = ee;
        if (from_out) from_out[i] = (uint8_t)f;
        // indices stay inside the read even if the path were malformed
        const KmerParam p = kmp[min(max(kk, 0), n_k - 1)];
        em_buf[lane] = emission(evp[min(max(ee, 0), n_e - 1)], p.mean, p.stdv, p.c);
      }
      if (t0 + 64 >= P) first_k = __shfl(kk, (P - 1 - t0) & 63);
      ck -= __shfl(pk, 63);
      ce -= __shfl(pe, 63);
      __syncthreads();
      if (lane == 0) {
        const int cnt = min(64, P - t0);
        for (int i = 0; i < cnt; ++i) sum_em += (double)em_buf[i];
      }
      __syncthreads();
    }
    if (lane == 0) {
      const double avg = sum_em / (double)P;
      const bool spanned = first_k == 0;  // last pair is always on k-mer n_k-1
      const bool fail = avg < -5.0 || !spanned || max_gap > 50;  // GPU/align_cuda.cl:134,139,873
      o.n_pairs[r] = fail ? 0 : P;
      if (o.path_len) o.path_len[r] = P;
      if (o.score) o.score[r] = best;
      if (o.avg_log_emission) o.avg_log_emission[r] = avg;
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint: d(sum_r w_r * score_r)/d(shift, scale, var, level_mean, level_stdv).
// Only D and U cells of the path carry an emission; per-cell terms are
// recomputed in double from the float inputs and reduced in double.
__global__ __launch_bounds__(256) void abea_adjoint_kernel(KernelArgs a, const float *grad_score,
                                                          float *grad_scal, double *grad_model,
                                                          int n_reads) {
  const int lane = threadIdx.x & 63;
  const int wave = uniform((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  for (int r = wave; r < n_reads; r += n_waves) {
    const ReadMeta m = load_meta(a.meta, r);
    const int P = (m.n_k > 0 && m.n_e > 0) ? a.out.path_len[r] : 0;
    const double w = (double)grad_score[r];
    const abea_scalings_t sc = a.scalings[r];
    const double scale = (double)sc.scale, shift = (double)sc.shift;
    const abea_aligned_pair_t *pairs = a.out.pairs + m.pair_off;
    const uint8_t *from = a.out.path_from + m.pair_off;
    const float *evp = a.ev + m.ev_off;
    const KmerParam *kmp = a.km + m.km_off;
    double gs = 0.0, gc = 0.0;
    for (int i = lane; i < P; i += 64) {
      if (from[i] == FROM_L) continue;
      const abea_aligned_pair_t pr = pairs[i];
      const int kk = min(max(pr.ref_pos, 0), m.n_k - 1);
      const int ee = min(max(pr.read_pos, 0), m.n_e - 1);
      const int rank = kmp[kk].rank & (ABEA_NUM_KMER - 1);
      const abea_model_t md = a.model[rank];
      const double mu = (double)md.level_mean, sg = (double)md.level_stdv;
      const double inv = 1.0 / sg;
      const double z = ((double)evp[ee] - (scale * mu + shift)) * inv;
      gs += z * inv;
      gc += z * mu * inv;
      atomicAdd(&grad_model[2 * rank], w * (z * scale * inv));
      atomicAdd(&grad_model[2 * rank + 1], w * ((z * z - 1.0) * inv));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      gs += __shfl_xor(gs, off);
      gc += __shfl_xor(gc, off);
    }
    if (lane == 0) {
      grad_scal[3 * r + 0] = (float)(w * gs);
      grad_scal[3 * r + 1] = (float)(w * gc);
      grad_scal[3 * r + 2] = 0.0f;
    }
  }
}

__global__ void abea_zero_f64_kernel(double *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// ---------------------------------------------------------------------------
hipError_t launch_pre(const KernelArgs &a, int q_begin, int q_count, hipStream_t s) {
  if (q_count <= 0) return hipSuccess;
  const int grid = q_count < 4096 ? q_count : 4096;
  hipLaunchKernelGGL(abea_pre_kernel, dim3(grid), dim3(256), 0, s, a, q_begin, q_count);
  return hipGetLastError();
}

hipError_t launch_core(const KernelArgs &a, int q_begin, int q_count, bool persistent,
                       int waves_per_block, int grid_blocks, hipStream_t s) {
  if (q_count <= 0) return hipSuccess;
  const int p = persistent ? 1 : 0;
  int grid = persistent ? grid_blocks : (q_count + waves_per_block - 1) / waves_per_block;
  if (grid < 1) grid = 1;
  switch (waves_per_block) {
    case 1:
      hipLaunchKernelGGL(abea_core_kernel<1>, dim3(grid), dim3(64), 0, s, a, q_begin, q_count, p);
      break;
    case 2:
      hipLaunchKernelGGL(abea_core_kernel<2>, dim3(grid), dim3(128), 0, s, a, q_begin, q_count, p);
      break;
    case 4:
      hipLaunchKernelGGL(abea_core_kernel<4>, dim3(grid), dim3(256), 0, s, a, q_begin, q_count, p);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int core_max_waves_per_block() { return 4; }

hipError_t launch_post(const KernelArgs &a, int q_begin, int q_count, int rows, hipStream_t s) {
  if (q_count <= 0) return hipSuccess;
  const int grid = q_count < 65536 ? q_count : 65536;
  switch (rows) {
    case 32:
      hipLaunchKernelGGL(abea_post_kernel<32>, dim3(grid), dim3(64), 0, s, a, q_begin, q_count);
      break;
    case 64:
      hipLaunchKernelGGL(abea_post_kernel<64>, dim3(grid), dim3(64), 0, s, a, q_begin, q_count);
      break;
    case 128:
      hipLaunchKernelGGL(abea_post_kernel<128>, dim3(grid), dim3(64), 0, s, a, q_begin, q_count);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_adjoint(const KernelArgs &a, const float *grad_score, float *grad_scalings,
                          double *grad_model, hipStream_t s) {
  const int n = a.n_queue;
  if (n <= 0) return hipSuccess;
  const int grid = (n + 3) / 4 < 16384 ? (n + 3) / 4 : 16384;
  hipLaunchKernelGGL(abea_adjoint_kernel, dim3(grid), dim3(256), 0, s, a, grad_score,
                     grad_scalings, grad_model, n);
  return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  const int grid = (int)(blocks < 1024 ? blocks : 1024);
  hipLaunchKernelGGL(abea_zero_f64_kernel, dim3(grid), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace abea
