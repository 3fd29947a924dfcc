// ABEA kernels for gfx950 (CDNA4, wave64).
//
//   core     per read, one wavefront: staging of the per-k-mer emission
//            parameters and SoA event means (replaces align_kernel_pre_2d,
//            GPU/align_cuda.cl:162-276), then the adaptive band walk
//            (replaces align_kernel_core_2d_shm, GPU/align_cuda.cl:297-609),
//            optionally followed by the read's traceback
//   post     best end cell, traceback over the trace, QC
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
// by one slot per band.  Slots outside the band window hold -inf, so
// out-of-band neighbours read -inf exactly as the reference's
// is_offset_valid() tests do.
//
// Trace layout: 2-bit from-codes (0 = D, 1 = U, 2 = L) in tiles of 8 bands;
// tile t holds one 32-bit word per lane, band 8t+i in bits [4i, 4i+4) as
// (code of slot 2l) | (code of slot 2l+1) << 2.  Cell (b, slot j) is bits
// 4(b&7) + 2(j&1) of word 64(b>>3) + (j>>1).
//
// Numerics follow f5c's CPU align(): float cells, emission in float with
// correctly rounded operations and no contraction (built with
// -ffp-contract=off), transition sums in double (the reference's constants are
// double, GPU/align_cuda.cl:355-359), argmax with the reference's tie order
// (GPU/align_cuda.cl:536-542).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "abea_internal.h"
#include "abea_wave.h"

namespace abea {

#define ABEA_NEG_INF (-__builtin_inff())

typedef __attribute__((address_space(4))) const float c_f32;
typedef __attribute__((address_space(4))) const int32_t c_i32;
typedef __attribute__((address_space(4))) const uint32_t c_u32;

constexpr uint32_t FROM_D = 0, FROM_U = 1, FROM_L = 2;  // GPU/align_cuda.cl:130-132
constexpr int kHalfW = ABEA_BANDWIDTH / 2;

// Uniform (scalar-cache) load of one read's metadata record.
__device__ __forceinline__ ReadMeta load_meta(const ReadMeta *meta, int r) {
  const c_i32 *p = (const c_i32 *)(meta + r);
  int32_t w[sizeof(ReadMeta) / 4];
#pragma unroll
  for (int i = 0; i < (int)(sizeof(ReadMeta) / 4); ++i) w[i] = p[i];
  ReadMeta m;
  __builtin_memcpy(&m, w, sizeof(m));
  return m;
}

// lane l receives lane (l - 1) mod 64: DPP wave_ror:1.
__device__ __forceinline__ float rotate_up(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13C, 0xF, 0xF, true));
}

__device__ __forceinline__ uint32_t lane_bits(float v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane);
}

// Signed-integer order key of a float's bits: for non-NaN values without -0,
// key(x) < key(y) exactly when x < y.  Band cells are never NaN or -0 (they
// are -inf or rounded sums), so the general band step compares them on the
// scalar unit through these keys, whatever their sign (the fast tiles use the
// shorter move_step, for reads whose cells are all <= 0).
__device__ __forceinline__ int32_t order_key(uint32_t u) {
  const int32_t i = (int32_t)u;
  return i ^ ((i >> 31) & 0x7fffffff);
}

constexpr uint32_t kNegInfBits = 0xff800000u;

// The band's two end slots, K (lower left) and K+99 (upper right), are one
// even and one odd slot: the even one sits in the parity-0 register at lane
// x0, the odd one in the parity-1 register at lane x1; np = 1 - (K & 1).
// Lane numbers are used mod 64 (the lane-select operand's low six bits).
struct EndLanes {
  int x0, x1, np;
};
__device__ __forceinline__ EndLanes end_lanes(int K) {
  const int p = K & 1, h = K >> 1;
  return EndLanes{p ? h + 50 : h, p ? h : h + 49, 1 - p};
}

// Suzuki's rule (GPU/align_cuda.cl:398-410) for a band whose cells are all
// <= 0 or -inf (no NaN, no -0): there x < y exactly when bits(x) > bits(y)
// as unsigned integers, -inf having the largest bits.  a0 / a1 are the bits
// of the end cells in the parity-0 / parity-1 registers; ll is a1 when K is
// odd.  On odd bands both ends at -inf move right: the upper-right bits are
// capped just below -inf's, so ll = -inf compares greater exactly when ur is
// -inf too or finite.  Returns the move (1 = right) and advances the lanes:
// a right move makes K's parity flip, x0 <- x1 + 1 and x1 <- x0 (with
// h = K >> 1: even K has (x0, x1) = (h, h+49), odd K (h+50, h)).
template <bool ODD>
__device__ __forceinline__ uint32_t move_step(EndLanes &el, uint32_t a0, uint32_t a1) {
  const int t = el.x1 + 1;
  uint32_t r, ll, ur;
  int x0n, x1n;
  if (ODD)
    asm("s_cmp_eq_u32 %[np], 0\n\t"
        "s_cselect_b32 %[ll], %[a1], %[a0]\n\t"
        "s_cselect_b32 %[ur], %[a0], %[a1]\n\t"
        "s_min_u32 %[ur], %[ur], 0xff7fffff\n\t"
        "s_cmp_gt_u32 %[ll], %[ur]\n\t"
        "s_cselect_b32 %[x0n], %[t], %[x0]\n\t"
        "s_cselect_b32 %[x1n], %[x0], %[x1]\n\t"
        "s_cselect_b32 %[r], 1, 0"
        : [r] "=&s"(r), [ll] "=&s"(ll), [ur] "=&s"(ur), [x0n] "=&s"(x0n), [x1n] "=&s"(x1n)
        : [np] "s"(el.np), [a0] "s"(a0), [a1] "s"(a1), [t] "s"(t), [x0] "s"(el.x0),
          [x1] "s"(el.x1)
        : "scc");
  else
    asm("s_cmp_eq_u32 %[np], 0\n\t"
        "s_cselect_b32 %[ll], %[a1], %[a0]\n\t"
        "s_cselect_b32 %[ur], %[a0], %[a1]\n\t"
        "s_cmp_gt_u32 %[ll], %[ur]\n\t"
        "s_cselect_b32 %[x0n], %[t], %[x0]\n\t"
        "s_cselect_b32 %[x1n], %[x0], %[x1]\n\t"
        "s_cselect_b32 %[r], 1, 0"
        : [r] "=&s"(r), [ll] "=&s"(ll), [ur] "=&s"(ur), [x0n] "=&s"(x0n), [x1n] "=&s"(x1n)
        : [np] "s"(el.np), [a0] "s"(a0), [a1] "s"(a1), [t] "s"(t), [x0] "s"(el.x0),
          [x1] "s"(el.x1)
        : "scc");
  el.x0 = x0n;
  el.x1 = x1n;
  el.np ^= (int)r;
  return r;
}

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int s) {
  return s ? ((x << s) | (x >> (64 - s))) : x;
}

// Lanes l (0..63) whose slot 2l+p lies in the circular slot range
// [first, first+span) (mod 128); span <= 100.
__device__ __forceinline__ uint64_t slot_lanes(int first, int span, int p) {
  if (span <= 0) return 0;
  const int t = (first - p) & 127;
  const int l0 = (t + 1) >> 1;
  const int cnt = ((t + span + 1) >> 1) - l0;
  return rotl64((1ull << cnt) - 1, l0 & 63);  // cnt <= 50
}

__device__ __forceinline__ bool lane_in(uint64_t mask) {
  return __builtin_amdgcn_inverse_ballot_w64(mask);
}

// f5c emission, GPU/align_cuda.cl:511-528: a = (x - m) / s;
// em = (-0.918938f - log_s) + (-0.5a)a, with c = -0.918938f - log_s staged
// per k-mer.  The quotient is the correctly rounded (x - m) / s, obtained from the
// correctly rounded reciprocal rs = 1/s by one fused residual correction
// (Markstein): q0 = n*rs, r = fma(-q0, s, n) (exact), a = fma(r, rs, q0).
// This equals IEEE division for the finite, non-underflowing quotients that
// occur here and costs 4 operations instead of the ~10 of a divide.
// The last step uses that halving is exact: RN(RN(-0.5a)*a) = -0.5*RN(a*a)
// whenever 0.5a^2 is a normal float, so c + (-0.5a)a = fma(RN(a*a), -0.5, c)
// with the same rounding.  For 0 < |a| < 2^-62 both sides round to c unless
// c is exactly 0 (level_log_stdv == -0.918938f).
__device__ __forceinline__ float emission(float x, float m, float s, float rs, float c) {
  const float n = x - m;
  const float q0 = n * rs;
  const float r = __builtin_fmaf(-q0, s, n);
  const float a = __builtin_fmaf(r, rs, q0);
  return __builtin_fmaf(a * a, -0.5f, c);
}

// One DP cell, GPU/align_cuda.cl:532-542 with double transition constants;
// the neighbour scores arrive widened to double (exact).  The reference keeps
// D unless U is strictly larger, then takes U when it equals the running
// maximum, and likewise for L: on the values that occur (no NaN, no -0) that
// is "later candidate wins ties", i.e. >=.  u / l report the two decisions.
// The score is the maximum of the three whichever wins a tie (equal floats
// have equal bits here), and L wins exactly when sl equals that maximum, so
// the cell costs one max3 and two compares.
__device__ __forceinline__ float dp_cell(double diag, double up, double left, float em,
                                         double lp_step, double lp_stay, double lp_skip,
                                         bool &u, bool &l) {
  const double emd = (double)em;
  const float sd = (float)((diag + lp_step) + emd);
  const float su = (float)((up + lp_stay) + emd);
  const float sl = (float)(left + lp_skip);
  u = su >= sd;
  const float m = __builtin_fmaxf(__builtin_fmaxf(sd, su), sl);
  l = sl == m;
  return m;
}

// lane `lane` of v takes the wave-uniform value x (v_writelane_b32)
extern "C" __device__ int abea_llvm_writelane(int x, int lane, int v) __asm(
    "llvm.amdgcn.writelane.i32");
__device__ __forceinline__ float put_lane(float v, float x, int lane) {
  return __builtin_bit_cast(float, abea_llvm_writelane(__builtin_bit_cast(int, x), lane,
                                                       __builtin_bit_cast(int, v)));
}

__device__ __forceinline__ double rotate_up_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x13C, 0xF, 0xF, true);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x13C, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Shift the four decision bits of one band into each lane's trace word:
// acc = acc * 16 + (u0 << 3 | l0 << 2 | u1 << 1 | l1), one add-with-carry per
// bit with the decision mask as carry-in.
__device__ __forceinline__ uint32_t shift_in4(uint32_t acc, uint64_t u0, uint64_t l0, uint64_t u1,
                                              uint64_t l1) {
  uint32_t out;
  uint64_t carry;
  asm("v_addc_co_u32 %0, %1, %2, %2, %3\n\t"
      "v_addc_co_u32 %0, %1, %0, %0, %4\n\t"
      "v_addc_co_u32 %0, %1, %0, %0, %5\n\t"
      "v_addc_co_u32 %0, %1, %0, %0, %6"
      : "=&v"(out), "=&s"(carry)
      : "v"(acc), "s"(u0), "s"(l0), "s"(u1), "s"(l1));
  return out;
}

// ---------------------------------------------------------------------------
// Staging of one read by the wave that walks it (in place of
// align_kernel_pre_2d, GPU/align_cuda.cl:162-276): per-k-mer emission
// parameters and model ranks, and the event means as a plain float array.
// Four k-mers (events) per lane and round, their loads issued (from clamped
// indices) before any of the round's stores.  The wave's later loads see the
// stores without a cache operation: a CU's vector L1 stays coherent with its
// own stores, and the walk reads the staged arrays through vector loads only.
//
// Returns whether every emission of the read is <= 0 (every staged
// -0.918938f - log_stdv is): then, with transition log-probabilities <= 0,
// every band cell is <= 0 or -inf, which the fast tiles' move decision
// relies on.  A read with a k-mer whose emission can be positive (a model
// stdv below 0.399) walks every band by the general band step.
__device__ __forceinline__ bool stage_read(const KernelArgs &a, const ReadMeta &m, int r,
                                           int lane) {
  const char *S = a.seq + a.seq_ptr[r];
  const abea_scalings_t sc = a.scalings[r];
  KmerParam *km = a.km + m.km_off;
  uint16_t *kr = a.km_rank + m.km_off;
  bool pos = false;  // a staged constant that is not <= 0 (positive, or NaN)
  for (int k0 = 0; k0 < m.n_k; k0 += 256) {
    uint32_t rank[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      uint32_t lo, hi;
      kmer_bytes(S, min(k0 + 64 * u + lane, m.n_k - 1), lo, hi);
      rank[u] = rank_lohi(lo, hi);
    }
    abea_model_t md[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) md[u] = a.model[rank[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + 64 * u + lane;
      if (k < m.n_k) {
        KmerParam p;
        p.mean = sc.scale * md[u].level_mean + sc.shift;  // two roundings, as the reference
        p.stdv = md[u].level_stdv;                        // scaling.var unused, GPU/align_cuda.cl:515
        p.rstdv = 1.0f / md[u].level_stdv;                // IEEE (correctly rounded) divide
        p.c = -0.918938f - md[u].level_log_stdv;          // cached log, GPU/align_cuda.cl:517-518
        pos = pos || !(p.c <= 0.0f);
        km[k] = p;
        kr[k] = (uint16_t)rank[u];
      }
    }
  }
  // the padding on both sides of the read's k-mer records and events
  for (int j = lane; j < kStagePad; j += 64) {
    const KmerParam pad{0.0f, 1.0f, 1.0f, -__builtin_inff()};
    km[-kStagePad + j] = pad;
    km[m.n_k + j] = pad;
    kr[-kStagePad + j] = 0;
    kr[m.n_k + j] = 0;
  }
  float *ev = a.ev + m.sev_off;
  for (int j = lane; j < kStagePad; j += 64) {
    ev[-kStagePad + j] = kPadEvent;
    ev[m.n_e + j] = kPadEvent;
  }
  for (int e0 = 0; e0 < m.n_e; e0 += 256) {
    float v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = m.ev_off + min(e0 + 64 * u + lane, m.n_e - 1);
      v[u] = a.events ? a.events[e].mean : a.ev_mean_in[e];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e0 + 64 * u + lane < m.n_e) ev[e0 + 64 * u + lane] = v[u];
  }
  wave_sync();
  return __builtin_amdgcn_ballot_w64(pos) == 0;
}

// The persistent core's queue head, cleared before each batched launch.
__global__ void abea_queue_reset_kernel(int *counter) {
  if (threadIdx.x == 0) *counter = 0;
}

// ---------------------------------------------------------------------------
// core: the band walk of one read by one wavefront.
//
// Per-lane state: the event and k-mer parameters under the lane's two slots,
// the scores of the last band (float, and widened to double for the next
// transition sums) and the diagonal sources one band further back.
struct Slots {
  float ev0, ev1;
  float km0m, km0s, km0r, km0c, km1m, km1s, km1r, km1c;
  float M0, M1;   // band b-1, parity 0 / 1
  double D0, D1;  // the same, widened
  double Lq;      // left neighbour of parity 0 at band b-1 = its diagonal at band b
  double Dq;      // D0 of band b-2 = diagonal of parity 1 at band b
  uint32_t acc;   // trace bits of the current 8-band tile
};

// Wave-uniform walk state.  K and E (lower-left k-mer and event of the last
// band) are kept as byte offsets of the next prefetch: koff = 16 (K + 128)
// into the staged k-mer records, eoff = 4 (E + 1) into the staged events.
struct Walk {
  uint32_t eoff, koff;
  float nx_ev, nk_m, nk_s, nk_r, nk_c;  // event E+1 and k-mer K+128
  uint32_t mv;                          // move bits of the current 32-band word
  float best;                           // end search so far: best end score
  int best_e;                           // and its event
};


// Event E+1 and k-mer K+128 as wave-uniform values.  Vector loads: the
// staged arrays are written in the same kernel, which rules out the scalar
// cache.
__device__ __forceinline__ float uniform_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ void prefetch(Walk &w, const float *evp, const KmerParam *kmp) {
  w.nx_ev = uniform_f32(*(const float *)((const char *)evp + w.eoff));
  const KmerParam k = *(const KmerParam *)((const char *)kmp + w.koff);
  w.nk_m = uniform_f32(k.mean);
  w.nk_s = uniform_f32(k.stdv);
  w.nk_r = uniform_f32(k.rstdv);
  w.nk_c = uniform_f32(k.c);
}

__device__ __forceinline__ int walk_k(const Walk &w) { return (int)(w.koff >> 4) - 128; }
__device__ __forceinline__ int walk_e(const Walk &w) { return (int)(w.eoff >> 2) - 1; }

// Events move up one slot per band; slot K_{b-1} (lane h, parity P) takes
// event E_{b-1}+1 whatever the move: on a down move it is the new bottom
// cell's event; on a right move that slot leaves the window and the value is
// exactly the event the next slot needs one band later.  On a right move the
// same slot becomes the window top, k-mer K_{b-1}+128.
template <int P>
__device__ __forceinline__ void shift_slots(Slots &s, const Walk &w, int h, bool right) {
  const float e_rot = rotate_up(s.ev1);
  if (P == 0) {
    s.ev1 = s.ev0;
    s.ev0 = put_lane(e_rot, w.nx_ev, h);
  } else {
    s.ev1 = put_lane(s.ev0, w.nx_ev, h);
    s.ev0 = e_rot;
  }
  if (right) {
    if (P == 0) {
      s.km0m = put_lane(s.km0m, w.nk_m, h);
      s.km0s = put_lane(s.km0s, w.nk_s, h);
      s.km0r = put_lane(s.km0r, w.nk_r, h);
      s.km0c = put_lane(s.km0c, w.nk_c, h);
    } else {
      s.km1m = put_lane(s.km1m, w.nk_m, h);
      s.km1s = put_lane(s.km1s, w.nk_s, h);
      s.km1r = put_lane(s.km1r, w.nk_r, h);
      s.km1c = put_lane(s.km1c, w.nk_c, h);
    }
  }
}

// The two cells of every lane for band b.
//   parity 0 (slot 2l):   up = D0, left = slot 2l-1 of b-1 = D1 of lane l-1,
//                         diag = that left one band earlier (Lq)
//   parity 1 (slot 2l+1): up = D1, left = D0, diag = D0 of b-2 (Dq)
struct Cells {
  float M0, M1;
  double left0;
  uint64_t u0, l0, u1, l1;
};
__device__ __forceinline__ Cells compute_cells(const Slots &s, double lp_step, double lp_stay,
                                               double lp_skip) {
  Cells c;
  c.left0 = rotate_up_f64(s.D1);
  const float em0 = emission(s.ev0, s.km0m, s.km0s, s.km0r, s.km0c);
  const float em1 = emission(s.ev1, s.km1m, s.km1s, s.km1r, s.km1c);
  bool u0, l0, u1, l1;
  c.M0 = dp_cell(s.Lq, s.D0, c.left0, em0, lp_step, lp_stay, lp_skip, u0, l0);
  c.M1 = dp_cell(s.Dq, s.D1, s.D0, em1, lp_step, lp_stay, lp_skip, u1, l1);
  c.u0 = __builtin_amdgcn_ballot_w64(u0);
  c.l0 = __builtin_amdgcn_ballot_w64(l0);
  c.u1 = __builtin_amdgcn_ballot_w64(u1);
  c.l1 = __builtin_amdgcn_ballot_w64(l1);
  return c;
}

__device__ __forceinline__ void commit_cells(Slots &s, const Cells &c, float M0, float M1,
                                             uint64_t u1, uint64_t l1) {
  s.acc = shift_in4(s.acc, c.u0, c.l0, u1, l1);
  s.Lq = c.left0;
  s.Dq = s.D0;
  s.M0 = M0;
  s.M1 = M1;
  s.D0 = (double)M0;
  s.D1 = (double)M1;
}

// Suzuki's rule on band b-1's lower-left (slot K) and upper-right (slot
// K+99) cells, GPU/align_cuda.cl:398-410: right iff ll < ur; both outside
// the matrix (-inf) -> right iff b is odd.  With K = 2h + P, slot K is lane h
// of parity P and slot K+99 is lane h+49+P of the other parity.
template <int P>
__device__ __forceinline__ void band_ends(const Slots &s, int h, uint32_t &llb, uint32_t &urb) {
  llb = lane_bits(P ? s.M1 : s.M0, h);
  urb = lane_bits(P ? s.M0 : s.M1, (h + 49 + P) & 63);
}

// One band of a fast tile (band_fast): branch-free; the scalar state is the
// two end lanes and the parity of K (EndLanes), plus K itself in edge tiles.
// Interior bands have every window cell inside the matrix, no trim cell and
// the last k-mer outside the window.  The events and k-mer parameters under
// the slots come from the tile's gathered operands (tile_fast).
// Masking: an in-window cell reads only in-window cells of the previous two
// bands or the two slots just outside the window (K-1 and K+100), so those
// two are set to -inf and the other out-of-window slots are left as they are
// (never read by an in-window cell).
//
// Per-read values an edge tile needs.
struct EdgeConst {
  int n_e, n_k, js;   // js: slot of the last k-mer (the end-search column)
  int jn;             // slot of k-mer n_k while it is in the tile's slot range, else -1
  double lp_trim;
};

template <bool ODD, bool EDGE>
__device__ __forceinline__ void band_fast(Slots &s, int &K, EndLanes &el, uint32_t &gm, int b,
                                          float ev0, float ev1, const KmerParam &k0,
                                          const KmerParam &k1, double lp_step, double lp_stay,
                                          double lp_skip, const EdgeConst &ec, float &best,
                                          int &best_e) {
  // Suzuki's rule (GPU/align_cuda.cl:398-410) on the scalar unit: the end
  // cells come out of the two registers at the lanes el holds, and the move
  // is decided on their bits (move_step).  gm collects the tile's moves,
  // first band in the highest bit; interior tiles recover K from them.
  const uint32_t r =
      move_step<ODD>(el, lane_bits(s.M0, el.x0 & 63), lane_bits(s.M1, el.x1 & 63));
  if (EDGE) K += (int)r;
  gm = (gm << 1) + r;
  Slots t = s;
  t.ev0 = ev0;
  t.ev1 = ev1;
  t.km0m = k0.mean;
  t.km0s = k0.stdv;
  t.km0r = k0.rstdv;
  t.km0c = k0.c;
  t.km1m = k1.mean;
  t.km1s = k1.stdv;
  t.km1r = k1.rstdv;
  t.km1c = k1.c;
  const Cells c = compute_cells(t, lp_step, lp_stay, lp_skip);
  // slots K-1 and K+100 of the new K: for even K, K+100 in the parity-0
  // register at lane x1 + 1 and K-1 in the parity-1 register at lane x0 - 1;
  // for odd K, K-1 at lane x1 and K+100 at lane x0
  const float ninf = ABEA_NEG_INF;
  float M0v = put_lane(c.M0, ninf, (el.x1 + el.np) & 63);
  float M1 = put_lane(c.M1, ninf, (el.x0 - el.np) & 63);
  uint64_t u1 = c.u1, l1 = c.l1;
  if (EDGE) {
    // cell (e, n_k) lies outside the matrix, but its skip transition reads
    // (e, n_k - 1) inside it and no emission: it is set to -inf, the value
    // a -inf skip would give (with the same decision bits, U and L, since
    // its other two candidates are -inf).  Cells beyond it then compute to
    // -inf from -inf neighbours and -inf emissions.
    if (ec.jn >= 0) {
      if (ec.jn & 1)
        M1 = put_lane(M1, ninf, ec.jn >> 1);
      else
        M0v = put_lane(M0v, ninf, ec.jn >> 1);
    }
    if (K < 0) {
      // the trim cell (b - 1, -1) in slot 127 (lane 63, parity 1), from-code
      // U, GPU/align_cuda.cl:432-443
      const int et = b - 1;
      const float tv = (et >= 0 && et < ec.n_e) ? (float)(ec.lp_trim * (double)(et + 1))
                                                 : ABEA_NEG_INF;
      M1 = put_lane(M1, tv, 63);
      u1 |= 1ull << 63;
      l1 &= ~(1ull << 63);
    }
    // the end search (GPU/align_cuda.cl:722-746) on the last k-mer's cell
    // (b - 1 - n_k, n_k - 1) while it is in the band's fill range
    const int os = (ec.n_k - 1) - K, e = b - 1 - ec.n_k;
    if (os >= 0 && os < ABEA_BANDWIDTH && e >= 0 && e <= ec.n_e - 1) {
      const float v = __builtin_bit_cast(float, lane_bits((ec.js & 1) ? M1 : M0v, ec.js >> 1));
      const float sc = (float)((double)v + (double)(ec.n_e - e) * ec.lp_trim);
      if (sc > best) {
        best = sc;
        best_e = e;
      }
    }
  }
  commit_cells(s, c, M0v, M1, u1, l1);
}

struct ReadConst {
  int n_e, n_k, js;
  double lp_stay, lp_step, lp_skip, lp_trim;
  const float *evp;      // the read's staged events
  const KmerParam *kmp;  // the read's staged k-mer records
  uint32_t *trace;  // the read's tiles (lane l's word of tile t at t * kTileWords + l)
  uint32_t *moves;
  int32_t *kck;
  c_u32 *replay;
};

// Any band: fill range from the matrix bounds (GPU/align_cuda.cl:457-470),
// the trim cell (e, -1) while K < 0 (GPU/align_cuda.cl:432-443), the end-cell
// column (k = n_k - 1) for the end search, and replayed moves.
template <int P>
__device__ __forceinline__ void band_general(Slots &s, Walk &w, const ReadConst &rc, int b,
                                             int lane) {
  int K = walk_k(w), E = walk_e(w);
  const int h = (K >> 1) & 63;
  uint32_t llb, urb;
  band_ends<P>(s, h, llb, urb);
  const bool both_out = llb == kNegInfBits && urb == kNegInfBits;
  bool right = both_out ? (b & 1) != 0 : order_key(llb) < order_key(urb);
  if (rc.replay) right = ((rc.replay[b >> 5] >> (b & 31)) & 1u) != 0;
  shift_slots<P>(s, w, h, right);
  if (right) {
    w.koff += 16;
    ++K;
  } else {
    w.eoff += 4;
    ++E;
  }
  w.mv |= (uint32_t)right << (b & 31);
  prefetch(w, rc.evp, rc.kmp);

  const Cells c = compute_cells(s, rc.lp_step, rc.lp_stay, rc.lp_skip);
  const int lo = max(max(-K, E - (rc.n_e - 1)), 0);
  const int hi = min(min(rc.n_k - K, E + 1), ABEA_BANDWIDTH);
  const int span = hi - lo;
  const int first = (K + lo) & 127;
  float M0 = lane_in(slot_lanes(first, span, 0)) ? c.M0 : ABEA_NEG_INF;
  float M1 = lane_in(slot_lanes(first, span, 1)) ? c.M1 : ABEA_NEG_INF;
  const bool trim = K < 0;
  if (trim) {
    // the trim cell (e, k=-1) sits in slot 127 (lane 63, parity 1)
    const int et = E - (-1 - K);
    const bool in = et >= 0 && et < rc.n_e;
    const float tv = in ? (float)(rc.lp_trim * (double)(et + 1)) : ABEA_NEG_INF;
    if (lane == 63) M1 = tv;
  }
  // End search (GPU/align_cuda.cl:722-746), folded into the walk: band b
  // holds column cell (e = b - n_k - 1, n_k - 1) when it is in the fill
  // range, so the cells come in increasing e and a running maximum with a
  // strict > keeps the reference's first maximum.
  const int os = (rc.n_k - 1) - K;
  if (span > 0 && (uint32_t)(os - lo) < (uint32_t)span) {
    const float v = __builtin_bit_cast(float, lane_bits((rc.js & 1) ? M1 : M0, rc.js >> 1));
    const int e = E - os;
    const float sc = (float)((double)v + (double)(rc.n_e - e) * rc.lp_trim);
    if (sc > w.best) {
      w.best = sc;
      w.best_e = e;
    }
  }
  commit_cells(s, c, M0, M1, c.u1, c.l1);
  // the trim cell's from-code is U: bits (u1, l1) = (1, 0) in lane 63
  if (trim && lane == 63) s.acc = (s.acc & ~3u) | 2u;
  if ((b & 7) == 7) {
    rc.trace[(b >> 3) * kTileWords + lane] = s.acc;
    s.acc = 0;
  }
  if ((b & 31) == 31) {
    rc.moves[b >> 5] = w.mv;  // same value from every lane
    rc.kck[b >> 5] = K;
    w.mv = 0;
  }
}

__device__ __forceinline__ void band_any(Slots &s, Walk &w, const ReadConst &rc, int b, int lane) {
  if (walk_k(w) & 1)
    band_general<1>(s, w, rc, b, lane);
  else
    band_general<0>(s, w, rc, b, lane);
}

// Operands of one 8-band tile, gathered from the staged arrays: the events
// under each lane's two slots on the 8 bands, and the two slots' k-mers.
struct TileOps {
  float e0[8], e1[8];
  KmerParam k0, k1;
};

// Bands b0 .. b0+7 (b0 a multiple of 8) when all of them are interior.
//
// Slot j holds k-mer K0 + ((j - K0) mod 128) for every k-mer the window can
// reach in the tile (K grows by at most 8, so the window stays inside
// [K0, K0+107]), and its event on band b is b - 2 - k.  So the tile's
// operands follow from K0 alone, and so do the next tile's: its window lies
// inside [K0, K0+115], where this tile's slot assignment still holds (the
// slots vacated here are below it).  Each tile therefore gathers the next
// one's operands into N (the first tile of a run gathers its own first,
// tile_gather).
//
// Index bounds: events b0 - 2 - K0 - off + i = E0 + 1 - off + i lie in
// [E0 - 126, E0 + 16] and k-mers K0 + off in [K0, K0 + 127].  Each read's
// staged arrays carry kStagePad = 128 padding entries on both sides, so fast
// tiles may start at -120 <= K0 <= n_k and 0 <= E0 <= n_e + 111.
// A scalar value the compiler takes as computed: keeps a uniform sum from
// being split into per-lane vector adds.
__device__ __forceinline__ uint32_t s_opaque(uint32_t v) {
  uint32_t r;
  asm("s_mov_b32 %0, %1" : "=s"(r) : "s"(v));
  return r;
}
struct TileAddr {
  const char *evq;    // the read's staged events, front padding included
  const char *kq;     // k-mer record K0
  uint32_t eo0, eo1;  // byte offset from evq of the event of slot 2l / 2l+1 on the tile's first band
  uint32_t ko0, ko1;  // byte offset from kq of the k-mer of slot 2l / 2l+1
};
// Byte offsets are taken from the start of the read's front padding, so that
// they stay non-negative: event b0 - 2 - K0 - off >= E0 - 126 >= -126.  The
// bases are wave-uniform and the offsets 32-bit, so the loads take a scalar
// base and a vector offset.
__device__ __forceinline__ TileAddr tile_addr(const ReadConst &rc, int b0, int K0, int lane) {
  // 4 (j - K0) mod 512 for slot j = 2l and 2l + 1
  const uint32_t q0 = (uint32_t)(8 * lane - 4 * K0) & 0x1fcu;
  const uint32_t q1 = (q0 + 4u) & 0x1fcu;
  const uint32_t eb4 = s_opaque(4u * (uint32_t)(b0 - 2 - K0 + kStagePad));  // >= q + 8
  TileAddr t;
  t.evq = (const char *)(rc.evp - kStagePad);
  t.kq = (const char *)(rc.kmp + K0);
  t.eo0 = eb4 - q0;
  t.eo1 = eb4 - q1;
  t.ko0 = 4u * q0;
  t.ko1 = 4u * q1;
  return t;
}
__device__ __forceinline__ const float *ev_at(const char *base, uint32_t off) {
  return (const float *)(base + off);
}
__device__ __forceinline__ KmerParam ld_km(const char *base, uint32_t off) {
  return *(const KmerParam *)(base + off);
}

// The first tile of a run gathers its own operands.
__device__ __forceinline__ void tile_gather(TileOps &T, const Walk &w, const ReadConst &rc, int b0,
                                            int lane) {
  const TileAddr t = tile_addr(rc, b0, walk_k(w), lane);
  const float *ep0 = ev_at(t.evq, t.eo0), *ep1 = ev_at(t.evq, t.eo1);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    T.e0[i] = ep0[i];
    T.e1[i] = ep1[i];
  }
  T.k0 = ld_km(t.kq, t.ko0);
  T.k1 = ld_km(t.kq, t.ko1);
}

// Edge tiles (EDGE) also cover bands whose window reaches outside the matrix
// on the k side: the trim column while K < 0, and k-mers n_k.. at the right
// edge.  Cells outside the matrix compute to -inf without masks: their
// emission is -inf (padding events and k-mers), and each one's left
// neighbour is such a cell too, except for slot n_k, whose left neighbour
// (e, n_k - 1) is in the matrix: while that slot holds k-mer n_k (K0 >=
// n_k - 127) its cell is set to -inf.  The trim cell is set in slot 127
// while K < 0, and the end search reads the last k-mer's column.  Cells
// outside the matrix on the e side (above the first event, below the last)
// compute to -inf the same way in both kinds of tile.
template <bool EDGE>
__device__ __forceinline__ void tile_fast(Slots &s, Walk &w, const ReadConst &rc, int b0,
                                          int lane, const TileOps &T, TileOps &N) {
  const int K0 = walk_k(w);
  int K = K0;
  uint32_t gm = 0;
  EndLanes el = end_lanes(K0);
  const TileAddr t = tile_addr(rc, b0, K0, lane);
  const float *ep0 = ev_at(t.evq, t.eo0), *ep1 = ev_at(t.evq, t.eo1);
  N.k0 = ld_km(t.kq, t.ko0);
  N.k1 = ld_km(t.kq, t.ko1);
  // slot n_k & 127 holds k-mer n_k through the tile iff K0 >= n_k - 127
  const EdgeConst ec{rc.n_e, rc.n_k, rc.js, K0 >= rc.n_k - 127 ? (rc.n_k & 127) : -1,
                     rc.lp_trim};
  float best = w.best;
  int best_e = w.best_e;
#pragma unroll
  for (int i = 0; i < 8; i += 2) {
    band_fast<false, EDGE>(s, K, el, gm, b0 + i, T.e0[i], T.e1[i], T.k0, T.k1, rc.lp_step,
                           rc.lp_stay, rc.lp_skip, ec, best, best_e);
    N.e0[i] = ep0[8 + i];
    N.e1[i] = ep1[8 + i];
    band_fast<true, EDGE>(s, K, el, gm, b0 + i + 1, T.e0[i + 1], T.e1[i + 1], T.k0, T.k1,
                          rc.lp_step, rc.lp_stay, rc.lp_skip, ec, best, best_e);
    N.e0[i + 1] = ep0[9 + i];
    N.e1[i + 1] = ep1[9 + i];
  }
  if (EDGE) {
    w.best = best;
    w.best_e = best_e;
  } else {
    K = K0 + __builtin_popcount(gm);
  }
  w.koff = 16u * (uint32_t)(K + 128);
  w.eoff = 4u * (uint32_t)(b0 + 6 - K);  // E + 1 = b - 1 - K after band b = b0+7
  w.mv |= (__builtin_bitreverse32(gm) >> 24) << (b0 & 31);  // gm: band b0 in bit 7
  (rc.trace + (b0 >> 3) * kTileWords)[(uint32_t)lane] = s.acc;
  s.acc = 0;
  if ((b0 & 31) == 24) {
    rc.moves[b0 >> 5] = w.mv;  // same value from every lane
    rc.kck[b0 >> 5] = K;
    w.mv = 0;
  }
}

// A run of fast tiles of one kind from band b (a multiple of 8 whose tile is
// of that kind); returns the band after the run.  Tiles come in pairs that
// trade the two operand buffers, so a tile's gathers for the next one land
// where that tile reads.  The walk's event index E never falls below its
// start (50), and K never decreases: an interior run ends for good once the window
// nears the last k-mer, an edge run at the start when the window has left
// the trim column.
template <bool EDGE>
__device__ __forceinline__ int fast_run(Slots &s, Walk &w, const ReadConst &rc, int b, int lane,
                                        int n_bands) {
  auto more = [&](int bb) {
    const int K0 = walk_k(w), E0 = walk_e(w);
    if (bb + 8 > n_bands || E0 > rc.n_e + 111) return false;
    const bool interior = K0 >= 0 && K0 + 8 < rc.n_k - ABEA_BANDWIDTH;
    return EDGE ? (K0 <= rc.n_k && !interior) : interior;
  };
  TileOps T0, T1;
  tile_gather(T0, w, rc, b, lane);
  do {
    tile_fast<EDGE>(s, w, rc, b, lane, T0, T1);
    b += 8;
    if (!more(b)) break;
    tile_fast<EDGE>(s, w, rc, b, lane, T1, T0);
    b += 8;
  } while (more(b));
  return b;
}

// Fast tiles leave the slot registers' events and k-mers, and the walk's
// scalar prefetches, behind; the general band expects what band b-1 left
// there: slot j with k-mer K + o and event E - o, o = (j - K) mod 128
// (E + K = b - 3).  After a fast tile E - o >= -127 and E <= n_e + 119 stay
// inside the read's event padding; k-mers up to K + 127 <= n_k + 135 may
// reach 8 entries past its k-mer padding, into the next read's front padding
// (or the slack after the last read), for slots outside the matrix whose
// cells the general band masks.
__device__ __forceinline__ void restore_slots(Slots &s, Walk &w, const ReadConst &rc,
                                              int lane) {
  const int K = walk_k(w), E = walk_e(w);
  const int o0 = (2 * lane - K) & 127, o1 = (2 * lane + 1 - K) & 127;
  s.ev0 = rc.evp[E - o0];
  s.ev1 = rc.evp[E - o1];
  const KmerParam p0 = rc.kmp[K + o0], p1 = rc.kmp[K + o1];
  s.km0m = p0.mean;
  s.km0s = p0.stdv;
  s.km0r = p0.rstdv;
  s.km0c = p0.c;
  s.km1m = p1.mean;
  s.km1s = p1.stdv;
  s.km1r = p1.rstdv;
  s.km1c = p1.c;
  prefetch(w, rc.evp, rc.kmp);  // event E+1 and k-mer K+128 for the next band
}

__device__ __forceinline__ void core_read(const KernelArgs &a, const ReadMeta &m, int r,
                                          int lane, bool nonpos) {
  ReadConst rc;
  rc.n_e = m.n_e;
  rc.n_k = m.n_k;
  rc.js = (m.n_k - 1) & 127;  // slot of the last k-mer (end-cell column)
  rc.lp_stay = m.lp_stay;
  rc.lp_step = m.lp_step;
  rc.lp_skip = a.lp_skip;
  rc.lp_trim = a.lp_trim;
  rc.evp = a.ev + m.sev_off;
  rc.kmp = a.km + m.km_off;
  rc.trace = a.trace + m.trace_off;
  rc.moves = a.moves + m.mv_off;
  rc.kck = a.kck + m.mv_off;
  rc.replay = a.replay ? (c_u32 *)(a.replay + m.mv_off) : nullptr;
  const int n_e = m.n_e, n_k = m.n_k;
  const int n_bands = n_e + n_k + 2;
  const float *evp = a.ev + m.sev_off;
  const KmerParam *kmp = a.km + m.km_off;

  // band 1: lower-left (E, K) = (50, -51), GPU/align_cuda.cl:251-255; slots
  // hold k-mers K..K+127 and the events on band 1's anti-diagonal
  const int E = kHalfW, K = -1 - kHalfW;
  Slots s;
  {
    const int o0 = (2 * lane - K) & 127, o1 = (2 * lane + 1 - K) & 127;
    const int e0 = E - o0, e1 = E - o1;
    s.ev0 = (e0 >= 0 && e0 < n_e) ? evp[e0] : 0.f;
    s.ev1 = (e1 >= 0 && e1 < n_e) ? evp[e1] : 0.f;
    const int k0 = K + o0, k1 = K + o1;
    KmerParam p0{0.f, 1.f, 1.f, 0.f}, p1{0.f, 1.f, 1.f, 0.f};
    if (k0 >= 0 && k0 < n_k) p0 = kmp[k0];
    if (k1 >= 0 && k1 < n_k) p1 = kmp[k1];
    s.km0m = p0.mean;
    s.km0s = p0.stdv;
    s.km0r = p0.rstdv;
    s.km0c = p0.c;
    s.km1m = p1.mean;
    s.km1s = p1.stdv;
    s.km1r = p1.rstdv;
    s.km1c = p1.c;
  }
  // band 1: only the first trim cell (e=0, k=-1 -> slot 127) = lp_trim
  // (GPU/align_cuda.cl:263-267); band 0: the start cell (e=-1, k=-1) = 0 in
  // slot 127, whose left-shifted copy is parity 0's diagonal at band 2
  s.M0 = ABEA_NEG_INF;
  s.M1 = (lane == 63) ? (float)rc.lp_trim : ABEA_NEG_INF;
  s.D0 = (double)s.M0;
  s.D1 = (double)s.M1;
  s.Lq = (lane == 0) ? 0.0 : (double)ABEA_NEG_INF;
  s.Dq = (double)ABEA_NEG_INF;
  s.acc = 0;
  Walk w;
  w.eoff = 4u * (uint32_t)(E + 1);
  w.koff = 16u * (uint32_t)(K + 128);
  w.mv = 0;
  w.best = ABEA_NEG_INF;
  w.best_e = 0;
  prefetch(w, rc.evp, rc.kmp);

  // Whole tiles of 8 bands take the fast path (over 8 bands K and E grow by
  // at most 8): interior tiles (kind 1) when the window stays clear of the
  // trim column and of the last k-mers, edge tiles (kind 2) otherwise.  The
  // gathers stay inside the staged padding for -120 <= K0 <= n_k and
  // 0 <= E0 <= n_e + 111.
  auto tile_kind = [&](int b) {
    const int K0 = walk_k(w), E0 = walk_e(w);
    if ((b & 7) != 0 || b + 8 > n_bands || rc.replay || !nonpos || K0 < -120 || K0 > n_k ||
        E0 < 0 || E0 > n_e + 111)
      return 0;
    return (K0 >= 0 && K0 + 8 < n_k - ABEA_BANDWIDTH) ? 1 : 2;
  };
  // The slot registers' events and k-mers are dead inside runs of fast
  // tiles and rebuilt when they end.  A read's runs go edge (start), interior,
  // edge (end), each with its own tile loop.
  for (int b = 2; b < n_bands;) {
    int kind = tile_kind(b);
    if (kind) {
      do {
        b = kind == 1 ? fast_run<false>(s, w, rc, b, lane, n_bands)
                      : fast_run<true>(s, w, rc, b, lane, n_bands);
        kind = tile_kind(b);
      } while (kind);
      if (b >= n_bands) break;
      restore_slots(s, w, rc, lane);
    }
    band_any(s, w, rc, b, lane);
    ++b;
  }
  const int last = n_bands - 1;
  if ((last & 7) != 7)
    rc.trace[(last >> 3) * kTileWords + lane] = s.acc << (4 * (7 - (last & 7)));
  if ((last & 31) != 31) rc.moves[last >> 5] = w.mv;
  a.ends[r] = EndCell{w.best, w.best_e};  // same value from every lane
}

// ---------------------------------------------------------------------------
// Traceback.

// From-code of cell (band b, slot j): the band's nibble is bits
// [4(7 - b%8), +4) of the lane's tile word, (u0 l0 u1 l1) from the top; a set
// l bit means L, else a set u bit means U, else D.
__device__ __forceinline__ uint32_t trace_code(const uint32_t *tw, int b, int j) {
  const uint32_t w = tw[(b >> 3) * kTileWords + (j >> 1)];
  const uint32_t bits = (w >> (4 * (7 - (b & 7)) + 2 * (1 - (j & 1)))) & 3u;
  return (bits & 1u) ? FROM_L : (bits >> 1);
}

// K (lower-left k-mer) of band b from the move bits and the checkpoints.
__device__ __forceinline__ int band_k(const uint32_t *mv, const int32_t *kck, int b) {
  if (b <= 1) return -1 - kHalfW;
  const int w = b >> 5, bit = b & 31;
  const int base = w ? kck[w - 1] : -1 - kHalfW;
  const uint32_t mask = bit == 31 ? 0xffffffffu : ((2u << bit) - 1u);
  return base + __builtin_popcount(mv[w] & mask);
}

// Traceback segments: C per lane (chains), 64 lanes.  The fused core walks
// one chain per lane: every chain starts OVERLAP rows above its segment, so
// fewer, longer segments cost fewer instructions, and the other waves of
// the SIMD hide the chain's load latency (36.5 against 37.8 ms per forward
// of the 16,384-read workload with four).  The separate post kernel, alone
// on the GPU, keeps four chains per lane to overlap their loads.
constexpr int kChainsFused = 1, kChainsPost = 4;

// Per-wave LDS record of a read's traceback segments: where the path
// following each segment's codes entered it (first cell at or below its top
// row), how many cells it recorded, and where it left (first cell below its
// bottom row; (-1, -1) when the path ended inside).
template <int C>
struct SegInfo {
  static constexpr int kMaxSeg = 64 * C;
  int ent_k[kMaxSeg], ent_e[kMaxSeg], cnt[kMaxSeg], ex_k[kMaxSeg], ex_e[kMaxSeg];
};

// Follow the from-codes of C segments at once: chain c starts at cell
// (k[c], e[c]) and records the codes of cells in rows [lo, hi] of segment
// sid[c] (one byte each, traceback order) at seg + sid*S, and its entry,
// count and exit in si.  The chains are independent, so their trace loads
// overlap; each keeps the last 8-byte trace chunk it read (4 slots of one
// 8-band tile), which the next steps usually reuse.
template <int C, class SI>
__device__ __forceinline__ void chase_chains(const uint32_t *tw, uint8_t *seg, int S, int b_end,
                                             const int (&sid)[C], int (&k)[C], int (&e)[C],
                                             SI *si) {
  // a segment's entry and exit cells go to its LDS record when they are
  // found (-1 until then), not held in registers
  int lo[C], hi[C], cnt[C], ct[C], cq[C];
  uint32_t word[C];
  uint2 cw[C];
  bool act[C], have[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    act[c] = sid[c] >= 0;
    lo[c] = 2 + sid[c] * S;
    hi[c] = min(lo[c] + S - 1, b_end);
    cnt[c] = 0;
    ct[c] = cq[c] = -1;
    word[c] = 0;
    cw[c] = make_uint2(0, 0);
    have[c] = false;
    if (act[c]) {
      si->ent_k[sid[c]] = -1;
      si->ent_e[sid[c]] = -1;
      si->ex_k[sid[c]] = -1;
      si->ex_e[sid[c]] = -1;
    }
  }
  for (;;) {
    bool any = false;
#pragma unroll
    for (int c = 0; c < C; ++c) any = any || act[c];
    if (__builtin_amdgcn_ballot_w64(any) == 0) break;
    int bb[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      bb[c] = e[c] + k[c] + 2;
      if (act[c]) {
        if (!have[c] && bb[c] <= hi[c]) {
          have[c] = true;
          si->ent_k[sid[c]] = k[c];
          si->ent_e[sid[c]] = e[c];
        }
        if (bb[c] < lo[c]) {
          si->ex_k[sid[c]] = k[c];
          si->ex_e[sid[c]] = e[c];
          act[c] = false;
        }
      }
      const int t = bb[c] >> 3, q = (k[c] & 127) >> 2;
      if (act[c] && (t != ct[c] || q != cq[c])) {
        ct[c] = t;
        cq[c] = q;
        cw[c] = *(const uint2 *)(tw + t * kTileWords + 2 * q);
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (act[c]) {
        const uint32_t wd = (k[c] & 2) ? cw[c].y : cw[c].x;
        const uint32_t bits = (wd >> (4 * (7 - (bb[c] & 7)) + 2 * (1 - (k[c] & 1)))) & 3u;
        const uint32_t f = (bits & 1u) ? FROM_L : (bits >> 1);
        if (bb[c] <= hi[c]) {
          word[c] |= f << (8 * (cnt[c] & 3));
          if ((cnt[c] & 3) == 3) {
            ((uint32_t *)(seg + sid[c] * S))[cnt[c] >> 2] = word[c];
            word[c] = 0;
          }
          ++cnt[c];
        }
        k[c] -= (f != FROM_U) ? 1 : 0;
        e[c] -= (f != FROM_L) ? 1 : 0;
        if (k[c] < 0 || e[c] < 0) act[c] = false;  // the path ended: exit stays (-1, -1)
      }
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    if (sid[c] >= 0) {
      if (cnt[c] & 3) ((uint32_t *)(seg + sid[c] * S))[cnt[c] >> 2] = word[c];
      si->cnt[sid[c]] = cnt[c];
    }
  }
}

// Follow the from-codes of one segment with the whole wave, from cell (k, e)
// (wave-uniform), recording as chase_chains does.  The trace comes in blocks
// of kBlockTiles 8-band tiles: each lane loads its word of every tile of the
// block at once (coalesced across the wave), and the path's codes are read
// out of the registers with v_readlane, so the walk through a block waits on
// no load.  The codes collect in a register, a word per lane, and are
// stored 256 at a time.  Where speculation failed this replaces one lane's
// chain, which waits on a load every few steps.
constexpr int kBlockTiles = 16;
template <class SI>
__device__ __forceinline__ void chase_wave(const uint32_t *tw, uint8_t *seg, int S, int b_end,
                                           int sid, int k, int e, SI *si, int lane) {
  const int lo = 2 + sid * S, hi = min(lo + S - 1, b_end);
  int ent_k = -1, ent_e = -1, ex_k = -1, ex_e = -1, cnt = 0;
  uint32_t word = 0, outv = 0;
  bool have = false, done = false;
  uint32_t *out = (uint32_t *)(seg + sid * S);
  int t0 = (e + k + 2) >> 3;  // the block's top tile
  while (!done) {
    uint32_t blk[kBlockTiles];
#pragma unroll
    for (int j = 0; j < kBlockTiles; ++j)
      blk[j] = t0 - j >= 0 ? tw[(t0 - j) * kTileWords + lane] : 0u;
    for (;;) {
      const int b = e + k + 2;
      if (!have && b <= hi) {
        have = true;
        ent_k = k;
        ent_e = e;
      }
      if (b < lo) {
        ex_k = k;
        ex_e = e;
        done = true;
        break;
      }
      const int j = t0 - (b >> 3);
      if (j >= kBlockTiles) break;  // the next block
      const uint32_t wd = (uint32_t)__builtin_amdgcn_readlane((int)blk[j], (k & 127) >> 1);
      const uint32_t bits = (wd >> (4 * (7 - (b & 7)) + 2 * (1 - (k & 1)))) & 3u;
      const uint32_t f = (bits & 1u) ? FROM_L : (bits >> 1);
      if (b <= hi) {
        word |= f << (8 * (cnt & 3));
        if ((cnt & 3) == 3) {
          const int wi = cnt >> 2;
          outv = lane == (wi & 63) ? word : outv;
          if ((wi & 63) == 63) out[wi - 63 + lane] = outv;
          word = 0;
        }
        ++cnt;
      }
      k -= (f != FROM_U) ? 1 : 0;
      e -= (f != FROM_L) ? 1 : 0;
      if (k < 0 || e < 0) {  // the path ended: exit stays (-1, -1)
        done = true;
        break;
      }
    }
    t0 -= kBlockTiles;
  }
  // the words not yet stored: those of the last, partial group of 64
  const int nw = cnt >> 2;  // complete words
  if (lane < (nw & 63)) out[(nw & ~63) + lane] = outv;
  if (lane == 0) {
    if (cnt & 3) out[nw] = word;
    si->ent_k[sid] = ent_k;
    si->ent_e[sid] = ent_e;
    si->ex_k[sid] = ex_k;
    si->ex_e[sid] = ex_e;
    si->cnt[sid] = cnt;
  }
}

// Traceback of one read by one wavefront (the body of the post kernel, and
// the tail of the fused core).
//  1. best end cell on the last k-mer (GPU/align_cuda.cl:722-746), wave-wide
//     with the reference's first-maximum rule.
//  2. traceback (GPU/align_cuda.cl:755-828), segment-parallel: the rows the
//     path spans are cut into 64 segments, one per lane.  Each lane follows
//     the from-codes through its segment starting OVERLAP rows above it (from
//     the band centre there, or from the true end cell when that is within
//     reach).  Tracebacks from different cells merge, so the lane is on the
//     true path by the time it enters its segment.  This is then checked
//     exactly, top down: a segment is accepted only if the cell where the lane
//     entered it is the cell where the true path (the accepted segment above)
//     leaves for it; otherwise that segment is followed again from the true
//     cell.  The result is the reference's path, whatever the speculation.
//  3. pairs in ascending order (no reversal pass), rebuilt from the end cell
//     with a wave prefix sum over the moves; the QC emission sum runs in the
//     reference's (traceback) order, in double; max gap from ballots.
//  4. QC (GPU/align_cuda.cl:856-882).
template <int C>
__device__ void traceback_read(const KernelArgs &a, const ReadMeta &m, int r, int lane,
                               int overlap, float *em_buf, SegInfo<C> *si) {
  constexpr int kMaxSeg = SegInfo<C>::kMaxSeg;
  const int n_e = m.n_e, n_k = m.n_k;
  const abea_output_t &o = a.out;
  // the end search ran with the band walk (band_general)
  float best = ABEA_NEG_INF;
  int best_e = 0;
  if (n_k > 0 && n_e > 0) {
    // a vector load: in the fused kernel this wave has just stored it, and the
    // scalar cache does not see vector stores
    const EndCell ec = a.ends[r];
    best = __builtin_bit_cast(float, uniform(__builtin_bit_cast(int, ec.score)));
    best_e = uniform(ec.e);
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
    return;
  }

  const uint32_t *tw = a.trace + m.trace_off;
  const uint32_t *mvw = a.moves + m.mv_off;
  const int32_t *kck = a.kck + m.mv_off;
  uint8_t *seg = a.seg + m.seg_off;
  const int e_end = uniform(best_e), k_end = n_k - 1;
  const int b_end = e_end + k_end + 2;
  // rows 2..b_end in up to 64 C segments of S rows (S a multiple of 4:
  // word-aligned code storage per segment), C consecutive ones per lane
  const int rows = b_end - 1;
  const int S = max(4, (((rows + kMaxSeg - 1) / kMaxSeg) + 3) & ~3);
  const int nseg = (rows + S - 1) / S;

  // ---- phase A: every lane follows its segments, starting overlap rows
  // above each (from the band centre there), or at the end cell
  {
    int sid[C], k[C], e[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = lane * C + c;
      sid[c] = i < nseg ? i : -1;
      k[c] = k_end;
      e[c] = e_end;
      const int hi = min(2 + i * S + S - 1, b_end);
      const int s = hi + overlap;
      if (i < nseg && s < b_end) {
        const int Ks = band_k(mvw, kck, s), Es = s - 2 - Ks;
        const int slo = max(max(-Ks, Es - (n_e - 1)), 0);
        const int shi = min(min(n_k - Ks, Es + 1), ABEA_BANDWIDTH);
        if (shi > slo) {
          const int off = min(max(kHalfW, slo), shi - 1);
          k[c] = Ks + off;
          e[c] = Es - off;
        }
      }
    }
    chase_chains<C>(tw, seg, S, b_end, sid, k, e, si);
  }
  wave_sync();

  // ---- phase B: accept segments top down, following again where needed
  int tk = k_end, te = e_end;
  int P = 0;
  for (int i = nseg - 1; i >= 0; --i) {
    if (tk < 0) {  // the path ended above this segment
      si->cnt[i] = 0;
      continue;
    }
    if (si->ent_k[i] != tk || si->ent_e[i] != te) {
      chase_wave(tw, seg, S, b_end, i, tk, te, si, lane);
      wave_sync();
    }
    P += si->cnt[i];
    tk = si->ex_k[i];
    te = si->ex_e[i];
  }
  wave_sync();  // segment codes and counts are visible to the whole wave

  // ---- phase C: ascending pairs, QC statistics.  Codes are taken 256 at a
  // time (4 runs of 64 consecutive ones) so that their gathers overlap.
  const float *evp = a.ev + m.sev_off;
  const KmerParam *kmp = a.km + m.km_off;
  abea_aligned_pair_t *pairs = o.pairs + m.pair_off;
  uint8_t *from_out = o.path_from ? o.path_from + m.pair_off : nullptr;
  int ck = k_end, ce = e_end;  // position of the next pair in traceback order
  // The reference adds the emissions one by one, in double, in traceback
  // order.  While every partial sum is exact the order is immaterial, and a
  // wave reduction gives the same bits: that holds when each term is 0 or
  // at least 2^-7 in magnitude (a multiple of 2^-30) and the partial sums
  // stay below 2^23.  Otherwise the terms are added serially, in order.
  double sum_em = 0.0;
  bool exact = true;
  int first_k = k_end, run = 0, max_gap = 0;
  int t_base = 0;  // traceback index of the current segment's first code
  for (int i = nseg - 1; i >= 0; --i) {
    const int ci = si->cnt[i];
    const uint8_t *sc = seg + i * S;
    for (int u0 = 0; u0 < ci; u0 += 256) {
      uint32_t f[4];
      int kk[4], ee[4];
      bool act[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int u = u0 + 64 * j + lane;
        act[j] = u < ci;
        f[j] = act[j] ? sc[u] : FROM_D;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // move from ascending pair i-1 to pair i (= inverse of the from code):
        // k steps in the low half, e steps in the high half, scanned together
        const int d = (act[j] && f[j] != FROM_U ? 1 : 0) | (act[j] && f[j] != FROM_L ? 0x10000 : 0);
        const int pd = wave_scan_add(d) - d;  // exclusive
        kk[j] = ck - (pd & 0xffff);
        ee[j] = ce - (pd >> 16);
        const int tot = __builtin_amdgcn_readlane(pd + d, 63);
        ck -= tot & 0xffff;
        ce -= tot >> 16;
      }
      float em[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // indices stay inside the read even if the path were malformed
        const KmerParam p = kmp[min(max(kk[j], 0), n_k - 1)];
        em[j] = emission(evp[min(max(ee[j], 0), n_e - 1)], p.mean, p.stdv, p.rstdv, p.c);
        if (act[j]) {
          const int idx = P - 1 - (t_base + u0 + 64 * j + lane);
          pairs[idx].ref_pos = kk[j];
          pairs[idx].read_pos = ee[j];
          if (from_out) from_out[idx] = (uint8_t)f[j];
        } else {
          em[j] = 0.0f;
        }
      }
      const int nall = min(256, ci - u0);
      {
        const int jl = (nall - 1) >> 6;  // run holding the chunk's last code
        const int kl = jl == 0 ? kk[0] : jl == 1 ? kk[1] : jl == 2 ? kk[2] : kk[3];
        first_k = __builtin_amdgcn_readlane(kl, (nall - 1) & 63);
      }
      // max run of consecutive L codes (the reference's gap counter), run by run
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int nact = min(64, nall - 64 * j);
        if (nact <= 0) break;
        const uint64_t full = nact == 64 ? ~0ull : ((1ull << nact) - 1);
        const uint64_t lm = __ballot(act[j] && f[j] == FROM_L) & full;
        if (lm == full) {
          run += nact;
          max_gap = max(max_gap, run);
        } else {
          const int lead = __builtin_ctzll(~lm);
          max_gap = max(max_gap, run + lead);
          uint64_t x = lm;
          int rmax = 0;
          while (x) {
            x &= x >> 1;
            ++rmax;
          }
          max_gap = max(max_gap, rmax);
          const uint64_t zeros = ~lm & full;
          run = nact - 1 - (63 - __builtin_clzll(zeros));
        }
      }
      // emission sum
      bool grid_ok = true;
      double part = 0.0;
      float amax = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t bits = __builtin_bit_cast(uint32_t, em[j]);
        const uint32_t ex = (bits >> 23) & 0xffu;  // inf and NaN take the serial path
        grid_ok = grid_ok && ((bits & 0x7fffffffu) == 0 || (ex >= 120u && ex != 255u));
        part += (double)em[j];
        amax = __builtin_fmaxf(amax, __builtin_fabsf(em[j]));
      }
      // every partial sum is bounded by |sum| + 256 * max |term|
      const float mag = wave_max_nonneg(amax);
      exact = exact && __ballot(!grid_ok) == 0 &&
              __builtin_fabs(sum_em) + 256.0 * (double)mag < 8388608.0;
      if (exact) {
        sum_em += wave_sum_exact_f64(part);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) em_buf[64 * j + lane] = em[j];
        wave_sync();
        double acc = sum_em;
        if (lane == 0) {
          for (int q = 0; q < nall; ++q) acc += (double)em_buf[q];
        }
        const uint64_t ab = __builtin_bit_cast(uint64_t, acc);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ab);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(ab >> 32));
        sum_em = __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
        wave_sync();
      }
    }
    t_base += ci;
  }
  if (lane == 0) {
    const double avg = sum_em / (double)P;
    const bool spanned = first_k == 0;  // the last pair is always on k-mer n_k-1
    const bool fail = avg < -5.0 || !spanned || max_gap > 50;  // GPU/align_cuda.cl:134,139,873
    o.n_pairs[r] = fail ? 0 : P;
    if (o.path_len) o.path_len[r] = P;
    if (o.score) o.score[r] = best;
    if (o.avg_log_emission) o.avg_log_emission[r] = avg;
  }
}

// Persistent (or one-read-per-wave) core; FUSE appends the traceback of each
// read to its band walk in the same wave.  Five waves per SIMD (at most 96
// VGPRs): the band walk fits, and the fused traceback spills a few values
// to scratch; the fifth wave hides more of the walk's latencies.  This
// measured 39.9 -> 38.6 ms on the 16,384-read workload.
template <int WPB, bool FUSE>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(5))) void abea_core_kernel(KernelArgs a, int q_begin, int q_count,
                                                             int persistent, int overlap) {
  __shared__ float em_bufs[WPB][256];
  __shared__ SegInfo<kChainsFused> seg_infos[FUSE ? WPB : 1];
  const int lane = threadIdx.x & 63;
  const int wib = uniform(threadIdx.x >> 6);
  int qi;
  if (persistent) {
    int t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1);
    qi = uniform(t);
  } else {
    qi = uniform(blockIdx.x * WPB + wib);
  }
  while (qi < q_count) {
    const int r = ((c_i32 *)a.queue)[q_begin + qi];
    const ReadMeta m = load_meta(a.meta, r);
    // The longest reads (queued first) set the kernel's critical path: one
    // wave walks a read band after band.  Their waves take issue priority
    // over the others on their SIMD.
    if (q_begin + qi < a.n_prio)
      __builtin_amdgcn_s_setprio(3);
    else
      __builtin_amdgcn_s_setprio(0);
    if (m.n_k > 0 && m.n_e > 0) {
      const bool nonpos = stage_read(a, m, r, lane);
      core_read(a, m, r, lane, nonpos);
    }
    if constexpr (FUSE) {
      wave_sync();  // this wave's trace, moves and end column are visible to it
      traceback_read(a, m, r, lane, overlap, em_bufs[wib], &seg_infos[wib]);
    }
    if (!persistent) break;
    int t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1);
    qi = uniform(t);
  }
}

__global__ __launch_bounds__(64) void abea_post_kernel(KernelArgs a, int q_begin, int q_count,
                                                      int overlap) {
  __shared__ float em_buf[256];
  __shared__ SegInfo<kChainsPost> seg_info;
  const int lane = threadIdx.x;
  for (int qi = blockIdx.x; qi < q_count; qi += gridDim.x) {
    const int r = ((c_i32 *)a.queue)[q_begin + qi];
    const ReadMeta m = load_meta(a.meta, r);
    traceback_read(a, m, r, lane, overlap, em_buf, &seg_info);
  }
}

// ---------------------------------------------------------------------------
// adjoint: d(sum_r w_r * score_r)/d(shift, scale, var) per read and
// d/d(level_mean, level_stdv, level_log_stdv) per k-mer.  Only D and U cells
// of the path carry an emission; per-cell terms are recomputed in double from
// the float inputs.  With a = (x - (scale*mu + shift)) / sigma and
// em = c - lambda - a^2/2 (lambda = the cached level_log_stdv the forward
// reads, GPU/align_cuda.cl:517-518): d em/d mu = a*scale/sigma,
// d em/d sigma = a^2/sigma (lambda held), d em/d lambda = -1.
//
// Per-read sums: each lane adds the cells of its own contiguous chunks in
// path order, then a fixed butterfly, so the bits do not depend on timing.
//
// Per-k-mer sums: a lane adds a run of cells on one k-mer position in double,
// then adds the run's weighted term to the k-mer's bins as a fixed-point
// number (two 64-bit integer limbs, hi * 2^-12 + lo * 2^-44).  Integer
// addition is associative, so the total does not depend on the order the
// adds land in: two runs on the same inputs give the same bits.  Each term is
// quantised once, with an error below 2^-45; the low limb takes up to 2^31
// terms per k-mer without overflow.  Terms that are not finite, or of
// magnitude 2^40 or more (never for pA-scale inputs), go to a double
// accumulator in global memory instead; a sum of non-finite values does not
// depend on its order either.
//
// The bins are privatised: each workgroup (4 to 16 waves, persistent over
// reads; one per CU, since the bins fill most of a CU's LDS) keeps the bins of half the k-mer ranks in LDS (2048 ranks x 3 columns x
// 2 limbs x 8 B = 96 KB) and walks its reads' paths once per half, adding
// only the runs whose rank falls in that half with LDS atomics.  At the end
// of each half it stores its bins to its own slice of a partial-sum buffer;
// the finish kernel adds the workgroups' slices in block order.  The only
// global atomics left are those of the rare double terms.  The paths are
// short next to the work per cell, so walking them twice costs little.
constexpr int kAdjChunk = 8;                // path cells per lane and step
constexpr int kAdjMaxWaves = 16;            // waves per workgroup, at most
constexpr int kAdjPassRanks = ABEA_NUM_KMER / 2;
constexpr int kAdjPassBins = 3 * kAdjPassRanks;  // (rank, column) bins per pass
constexpr int kAdjBins = 3 * ABEA_NUM_KMER;
constexpr double kFixHi = 4096.0;           // 2^12
constexpr double kFixLo = 17592186044416.0; // 2^44: rest = lo * 2^-44, lo in [0, 2^32]

// Adds t to bin i (hi limb at bins[i], lo limb at bins[kAdjPassBins + i]).
__device__ __forceinline__ void fix_add_lds(unsigned long long *bins, double *special, int i,
                                            int gi, double t) {
  if (t == 0.0) return;
  if (!(fabs(t) < 1099511627776.0)) {  // non-finite or |t| >= 2^40
    atomicAdd(&special[gi], t);
    return;
  }
  const double f = floor(t * kFixHi);               // exact scaling, exact floor
  const double rest = t - f * (1.0 / kFixHi);       // in [0, 2^-12)
  const long long hi = (long long)f;
  const long long lo = (long long)rint(rest * kFixLo);  // in [0, 2^32]
  atomicAdd(&bins[i], (unsigned long long)hi);
  atomicAdd(&bins[kAdjPassBins + i], (unsigned long long)lo);
}

__global__ __launch_bounds__(64 * kAdjMaxWaves) void abea_adjoint_kernel(
    KernelArgs a, const float *grad_score, float *grad_scal, unsigned long long *partial,
    double *special, int n_reads) {
  __shared__ unsigned long long bins[2 * kAdjPassBins];
  const int lane = threadIdx.x & 63;
  const int wave = uniform(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  for (int pass = 0; pass < 2; ++pass) {
    const int rank_lo = pass * kAdjPassRanks;
    for (int i = threadIdx.x; i < 2 * kAdjPassBins; i += blockDim.x) bins[i] = 0ull;
    __syncthreads();
    for (int r = blockIdx.x * wpb + wave; r < n_reads; r += gridDim.x * wpb) {
      const ReadMeta m = load_meta(a.meta, r);
      const int P = (m.n_k > 0 && m.n_e > 0) ? a.out.path_len[r] : 0;
      const double w = (double)grad_score[r];
      const abea_scalings_t sc = a.scalings[r];
      const double scale = (double)sc.scale, shift = (double)sc.shift;
      const abea_aligned_pair_t *pairs = a.out.pairs + m.pair_off;
      const uint8_t *from = a.out.path_from + m.pair_off;
      const float *evp = a.ev + m.sev_off;
      const uint16_t *kr = a.km_rank + m.km_off;
      double gs = 0.0, gc = 0.0;
      for (int base = 0; base < P; base += 64 * kAdjChunk) {
        int cur_k = -1, cur_bin = -1;  // cur_bin < 0: the run's rank is outside this pass
        double s_mu = 0.0, s_sg = 0.0, s_n = 0.0;
        auto flush = [&]() {
          if (cur_bin >= 0) {
            const int gi = 3 * rank_lo + cur_bin;
            fix_add_lds(bins, special, cur_bin + 0, gi + 0, w * s_mu);
            fix_add_lds(bins, special, cur_bin + 1, gi + 1, w * s_sg);
            fix_add_lds(bins, special, cur_bin + 2, gi + 2, -w * s_n);
          }
        };
        // the lane's cells: all loads first (pairs and codes, then ranks, then
        // events and model rows), so that each stage's loads are in flight
        // together; indices are clamped into the read's arrays
        const int i0 = base + lane * kAdjChunk;
        int kk[kAdjChunk], ee[kAdjChunk], rk[kAdjChunk];
        bool use[kAdjChunk];
        float xv[kAdjChunk];
        abea_model_t md[kAdjChunk];
#pragma unroll
        for (int j = 0; j < kAdjChunk; ++j) {
          const int i = min(i0 + j, P - 1);
          const abea_aligned_pair_t pr = pairs[i];
          use[j] = i0 + j < P && from[i] != FROM_L;
          kk[j] = min(max(pr.ref_pos, 0), m.n_k - 1);
          ee[j] = min(max(pr.read_pos, 0), m.n_e - 1);
        }
#pragma unroll
        for (int j = 0; j < kAdjChunk; ++j) rk[j] = kr[kk[j]] & (ABEA_NUM_KMER - 1);
#pragma unroll
        for (int j = 0; j < kAdjChunk; ++j) {
          xv[j] = evp[ee[j]];
          md[j] = a.model[rk[j]];
        }
#pragma unroll
        for (int j = 0; j < kAdjChunk; ++j) {
          if (!use[j]) continue;
          if (kk[j] != cur_k) {
            flush();
            cur_k = kk[j];
            cur_bin = (unsigned)(rk[j] - rank_lo) < (unsigned)kAdjPassRanks ? 3 * (rk[j] - rank_lo)
                                                                             : -1;
            s_mu = s_sg = s_n = 0.0;
          }
          // the per-read terms need every cell, in the first pass only
          if (cur_bin < 0 && pass != 0) continue;
          const double mu = (double)md[j].level_mean, sg = (double)md[j].level_stdv;
          const double inv = 1.0 / sg;
          const double z = ((double)xv[j] - (scale * mu + shift)) * inv;
          gs += z * inv;
          gc += z * mu * inv;
          s_mu += z * scale * inv;
          s_sg += z * z * inv;
          s_n += 1.0;
        }
        flush();
      }
      if (pass == 0) {
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
    __syncthreads();
    // this workgroup's slice: hi limbs of all bins, then lo limbs
    unsigned long long *dst = partial + (size_t)blockIdx.x * 2 * kAdjBins + 3 * rank_lo;
    for (int i = threadIdx.x; i < kAdjPassBins; i += blockDim.x) {
      dst[i] = bins[i];
      dst[kAdjBins + i] = bins[kAdjPassBins + i];
    }
    __syncthreads();  // before the next pass clears the bins
  }
}

// Clears the double accumulator before an adjoint (a fill kernel rather than
// a memset, so the sequence can be captured in a graph).  The partial-sum
// slices need no clearing: every workgroup stores all of its bins.
__global__ void abea_adjoint_clear_kernel(double *special) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kAdjBins; i += gridDim.x * blockDim.x)
    special[i] = 0.0;
}

// grad_model[3 * rank + column] = hi * 2^-12 + lo * 2^-44 (+ the non-finite
// or huge terms, which then decide the value), the limbs summed over the
// workgroups' slices in block order.
__global__ void abea_adjoint_finish_kernel(const unsigned long long *partial, int n_blocks,
                                           const double *special, double *grad_model) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kAdjBins; i += gridDim.x * blockDim.x) {
    unsigned long long h = 0ull, l = 0ull;
    int b = 0;
    // eight slices' loads in flight at a time (integer sums: any order)
    for (; b + 8 <= n_blocks; b += 8) {
      unsigned long long hv[8], lv[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        hv[j] = partial[(size_t)(b + j) * 2 * kAdjBins + i];
        lv[j] = partial[(size_t)(b + j) * 2 * kAdjBins + kAdjBins + i];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        h += hv[j];
        l += lv[j];
      }
    }
    for (; b < n_blocks; ++b) {
      h += partial[(size_t)b * 2 * kAdjBins + i];
      l += partial[(size_t)b * 2 * kAdjBins + kAdjBins + i];
    }
    const double hi = (double)(long long)h;
    const double lo = (double)(long long)l;
    grad_model[i] = (hi * (1.0 / kFixHi) + lo * (1.0 / kFixLo)) + special[i];
  }
}

// refine: one Gauss-Newton step on each read's (scale, shift) along the path
// of the last forward.  With m = scale*mu + shift and a = (x - m)/sigma,
// every D or U cell contributes -a^2/2 to the score, so on a fixed path the
// score is a quadratic in (scale, shift) and one Newton step reaches its
// maximum: solve [Sum mu^2/s^2, Sum mu/s^2; Sum mu/s^2, Sum 1/s^2] d = g with
// g the path gradient (the adjoint's per-read terms).  Sums in double.
//
// With a start state (state non-null) the loop's first two steps also pick
// each read's start: step 0 follows the forward at the given scalings,
// keeps them, their score and their Newton step, and proposes the
// method-of-moments scalings instead; step 1 follows the forward at those
// and keeps whichever of the two scored higher, proceeding with its Newton
// step.  A read whose alignment is lost under its given scalings (its path
// wanders through the trim column and skips) carries no information about
// them along that path, so Newton steps from there do not recover it.
__global__ __launch_bounds__(256) void abea_refine_kernel(KernelArgs a, abea_scalings_t *scal_out,
                                                         float damping, int n_reads,
                                                         const int32_t *step,
                                                         const abea_scalings_t *start,
                                                         RefineState *state) {
  const int lane = threadIdx.x & 63;
  const int wave = uniform((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  const int mode = step ? min(uniform(*(volatile const int32_t *)step), 2) : 2;
  for (int r = wave; r < n_reads; r += n_waves) {
    const ReadMeta m = load_meta(a.meta, r);
    const int P = (m.n_k > 0 && m.n_e > 0) ? a.out.path_len[r] : 0;
    const abea_scalings_t sc = a.scalings[r];
    const double scale = (double)sc.scale, shift = (double)sc.shift;
    const abea_aligned_pair_t *pairs = a.out.pairs + m.pair_off;
    const uint8_t *from = a.out.path_from + m.pair_off;
    const float *evp = a.ev + m.sev_off;
    const uint16_t *kr = a.km_rank + m.km_off;
    double gs = 0.0, gc = 0.0, hss = 0.0, hsc = 0.0, hcc = 0.0;
    for (int i = lane; i < P; i += 64) {
      if (from[i] == FROM_L) continue;
      const abea_aligned_pair_t pr = pairs[i];
      const int kk = min(max(pr.ref_pos, 0), m.n_k - 1);
      const int ee = min(max(pr.read_pos, 0), m.n_e - 1);
      const abea_model_t md = a.model[kr[kk] & (ABEA_NUM_KMER - 1)];
      const double mu = (double)md.level_mean;
      const double inv = 1.0 / (double)md.level_stdv;
      const double z = ((double)evp[ee] - (scale * mu + shift)) * inv;
      const double i2 = inv * inv;
      gc += z * inv;
      gs += z * mu * inv;
      hcc += i2;
      hsc += mu * i2;
      hss += mu * mu * i2;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      gs += __shfl_xor(gs, off);
      gc += __shfl_xor(gc, off);
      hss += __shfl_xor(hss, off);
      hsc += __shfl_xor(hsc, off);
      hcc += __shfl_xor(hcc, off);
    }
    if (lane == 0) {
      abea_scalings_t o = sc;
      const double det = hss * hcc - hsc * hsc;
      if (P > 0 && det > 0.0) {
        const double d_scale = (hcc * gs - hsc * gc) / det;
        const double d_shift = (hss * gc - hsc * gs) / det;
        o.scale = (float)(scale + (double)damping * d_scale);
        o.shift = (float)(shift + (double)damping * d_shift);
      }
      if (mode == 0) {
        const float score = P > 0 ? a.out.score[r] : ABEA_NEG_INF;
        RefineState st;
        st.given = sc;
        st.given_step = make_float2(o.scale, o.shift);
        st.given_score = score;
        st.pad = 0;
        state[r] = st;
        const abea_scalings_t ms = start[r];
        if (__builtin_isfinite(ms.scale) && __builtin_isfinite(ms.shift)) {
          o = sc;
          o.scale = ms.scale;
          o.shift = ms.shift;
        }
      } else if (mode == 1) {
        const RefineState st = state[r];
        const float score = P > 0 ? a.out.score[r] : ABEA_NEG_INF;
        if (!(score > st.given_score)) {
          o = st.given;
          o.scale = st.given_step.x;
          o.shift = st.given_step.y;
        }
      }
      scal_out[r] = o;
    }
  }
}

// Advances the refinement loop's step counter (one thread; inside the
// captured step, after the refine kernel).
__global__ void abea_refine_advance_kernel(int32_t *step, int32_t set_to) {
  if (threadIdx.x == 0) *step = set_to >= 0 ? set_to : *step + 1;
}

// Host-buffer drop-ins: each read's pairs, from its slots (ReadMeta
// pair_off) to dst + coff[r], one wave per read, so that the device-to-host
// copy moves only the pairs.
__global__ __launch_bounds__(256) void abea_compact_pairs_kernel(const ReadMeta *meta,
                                                                const abea_aligned_pair_t *pairs,
                                                                const int32_t *n_pairs,
                                                                const int64_t *coff,
                                                                abea_aligned_pair_t *dst,
                                                                int n_reads) {
  const int lane = threadIdx.x & 63;
  const int wave = uniform((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  for (int r = wave; r < n_reads; r += n_waves) {
    const int np = n_pairs[r];
    const abea_aligned_pair_t *src = pairs + meta[r].pair_off;
    abea_aligned_pair_t *d = dst + coff[r];
    for (int i = lane; i < np; i += 64) d[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
hipError_t launch_queue_reset(const KernelArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(abea_queue_reset_kernel, dim3(1), dim3(64), 0, s, a.counter);
  return hipGetLastError();
}

hipError_t launch_compact_pairs(const ReadMeta *meta, const abea_aligned_pair_t *pairs,
                                const int32_t *n_pairs, const int64_t *coff,
                                abea_aligned_pair_t *dst, int n_reads, hipStream_t s) {
  if (n_reads <= 0) return hipSuccess;
  const int grid = std::min((n_reads + 3) / 4, 4096);
  hipLaunchKernelGGL(abea_compact_pairs_kernel, dim3(grid), dim3(256), 0, s, meta, pairs, n_pairs,
                     coff, dst, n_reads);
  return hipGetLastError();
}

template <int WPB, bool FUSE>
static void launch_core_t(const KernelArgs &a, int q_begin, int q_count, int p, int grid,
                          int overlap, hipStream_t s) {
  hipLaunchKernelGGL((abea_core_kernel<WPB, FUSE>), dim3(grid), dim3(64 * WPB), 0, s, a, q_begin,
                     q_count, p, overlap);
}

hipError_t launch_core(const KernelArgs &a, int q_begin, int q_count, bool persistent,
                       int waves_per_block, int grid_blocks, int fuse, int overlap, hipStream_t s) {
  if (q_count <= 0) return hipSuccess;
  const int p = persistent ? 1 : 0;
  int grid = persistent ? grid_blocks : (q_count + waves_per_block - 1) / waves_per_block;
  if (grid < 1) grid = 1;
#define ABEA_CORE_CASE(W_, F_)                                        \
  if (waves_per_block == W_ && (fuse != 0) == F_) {                  \
    launch_core_t<W_, F_>(a, q_begin, q_count, p, grid, overlap, s); \
    return hipGetLastError();                                        \
  }
  ABEA_CORE_CASE(1, false)
  ABEA_CORE_CASE(2, false)
  ABEA_CORE_CASE(4, false)
  ABEA_CORE_CASE(1, true)
  ABEA_CORE_CASE(2, true)
  ABEA_CORE_CASE(4, true)
#undef ABEA_CORE_CASE
  return hipErrorInvalidValue;
}

int core_max_waves_per_block() { return 4; }

hipError_t launch_post(const KernelArgs &a, int q_begin, int q_count, int overlap, hipStream_t s) {
  if (q_count <= 0) return hipSuccess;
  const int grid = q_count < 65536 ? q_count : 65536;
  hipLaunchKernelGGL(abea_post_kernel, dim3(grid), dim3(64), 0, s, a, q_begin, q_count, overlap);
  return hipGetLastError();
}

// One workgroup per CU when there are reads for every CU; its waves: as many
// as the reads give each CU, between 4 and 16.
int adjoint_blocks(int n_reads, int n_cu) {
  const int want = (n_reads + 3) / 4;
  return std::max(1, std::min(want, n_cu));
}

hipError_t launch_adjoint(const KernelArgs &a, const float *grad_score, float *grad_scalings,
                          double *grad_model, unsigned long long *partial, int n_blocks,
                          double *special, hipStream_t s) {
  const int n = a.n_queue;
  const int wpb = std::max(4, std::min(kAdjMaxWaves, (n + n_blocks - 1) / std::max(n_blocks, 1)));
  hipLaunchKernelGGL(abea_adjoint_clear_kernel, dim3(48), dim3(256), 0, s, special);
  hipLaunchKernelGGL(abea_adjoint_kernel, dim3(n_blocks), dim3(64 * wpb), 0, s, a, grad_score,
                     grad_scalings, partial, special, n);
  // one 64-thread block per 64 bins: the sums spread over 192 CUs
  hipLaunchKernelGGL(abea_adjoint_finish_kernel, dim3(kAdjBins / 64), dim3(64), 0, s, partial,
                     n_blocks, special, grad_model);
  return hipGetLastError();
}

hipError_t launch_refine(const KernelArgs &a, abea_scalings_t *scal_out, float damping,
                         const int32_t *step, const abea_scalings_t *start, RefineState *state,
                         hipStream_t s) {
  const int n = a.n_queue;
  if (n <= 0) return hipSuccess;
  const int grid = (n + 3) / 4 < 16384 ? (n + 3) / 4 : 16384;
  hipLaunchKernelGGL(abea_refine_kernel, dim3(grid), dim3(256), 0, s, a, scal_out, damping, n,
                     step, start, state);
  return hipGetLastError();
}

hipError_t launch_refine_step_counter(int32_t *step, int32_t set_to, hipStream_t s) {
  hipLaunchKernelGGL(abea_refine_advance_kernel, dim3(1), dim3(64), 0, s, step, set_to);
  return hipGetLastError();
}

}  // namespace abea
