// Event detection for gfx950: f5c's getevents() (GPU/f5cmisc.h:44), the
// step before estimate_scalings_using_mom() and align().  The reference
// declares it and its types (event_t / event_table, adapted from scrappie,
// GPU/f5c.h:112-130); the body is f5c's port of scrappie's detector:
// prefix sums of the signal and its squares (double), two windowed
// t-statistics (w = 3 and 6), the short/long peak detector (thresholds 1.4
// and 9.0, peak height 0.2), and events between consecutive boundaries.
//
// Every stage runs over chunks of every read at once:
//   chunk_sum / read_scan   the double sums are exact in any order when
//            every term is a multiple of 2^g and the sum of magnitudes stays
//            below 2^(52+g) (true for pA-scale signals).  Then every
//            difference of prefix sums the C source takes (window sums,
//            event sums) is the exact sum of its samples, whatever the
//            order: those sums come straight from the samples, and only the
//            chunks' carries are kept (for long events).  A read that fails
//            the test takes one serial pass in index order that writes the
//            prefix sums, and its statistics and events read them.
//   detect   one lane per 512-sample chunk, started fresh a few samples
//            early (speculation): the chunk's entry state, its boundaries in
//            emission order and its exit state.  The lane computes the two
//            t-statistics of each sample as it goes (the C source's
//            float/double types): from window sums that slide along the
//            samples (exact reads), or from the prefix sums.
//   check / fix   a chunk is right when its entry state is its
//            predecessor's exit state (the detector is a deterministic state
//            machine and the first chunk starts from the true initial
//            state); from a read's first chunk where that fails, one lane
//            walks on, re-running each chunk whose entry is not the true
//            state.  The result is the serial one.
//   scan / events   per detector chunk, the index of its first event and the
//            boundary before it; one thread per event.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "abea_internal.h"
#include "abea_wave.h"

namespace abea {

namespace {

constexpr float kThrShort = 1.4f, kThrLong = 9.0f, kPeakHeight = 0.2f;
constexpr int kWShort = 3, kWLong = 6;

// --------------------------------------------------------------------------
// prefix sums

__device__ __forceinline__ double wave_scan_add_f64(double v) {
  v += dpp_f64(v, 0x111, 0xf);
  v += dpp_f64(v, 0x112, 0xf);
  v += dpp_f64(v, 0x114, 0xf);
  v += dpp_f64(v, 0x118, 0xf);
  v += dpp_f64(v, 0x142, 0xa);
  v += dpp_f64(v, 0x143, 0xc);
  return v;
}

__device__ __forceinline__ int wave_min_i32(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum_any_f64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// wave reductions over the exactness data of 64 values
__device__ __forceinline__ void reduce_sum_info(GeChunkSum &c) {
  c.s = wave_sum_any_f64(c.s);
  c.q = wave_sum_any_f64(c.q);
  c.abs_s = wave_sum_any_f64(c.abs_s);
  c.abs_q = wave_sum_any_f64(c.abs_q);
  c.gx = wave_min_i32(c.gx);
  c.gq = wave_min_i32(c.gq);
  c.bad = wave_max_i32(c.bad);
  c.hx = wave_max_i32(c.hx);
}

// ulp exponent of the float whose magnitude has bits u (-149 for
// subnormals); 1000 for zero (no constraint), and 2000 for inf/NaN (never
// exact)
__device__ __forceinline__ int ulp_exp_bits(uint32_t u) {
  if (u == 0) return 1000;
  const int e = (int)(u >> 23);
  if (e >= 255) return 2000;
  return e == 0 ? -149 : e - 150;
}

// One wave per prefix chunk (4 per block): the chunk's sums and exactness
// data, read 64 consecutive samples per step.  The ulp exponent of a finite
// float grows with the bits of its magnitude, so each lane keeps the
// smallest and largest magnitude bits of its samples and of their squares
// (zero counted as the largest unsigned value in the minimum) and converts
// them once: the smallest ulp exponent of the non-zero values, the largest
// one, and whether an inf or NaN (magnitude bits >= 0x7f800000) occurred.
// The squares are non-negative, so their sum of magnitudes is their sum.
__global__ __launch_bounds__(256) void abea_ge_chunk_sum_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= a.n_chunks) return;
  const int r = a.chunk_read[c];
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const int64_t c0 = (c - a.chunk_off[r]) * kGeChunk, c1 = min(c0 + kGeChunk, n);
  const float *x = a.raw + off;
  double s = 0.0, q = 0.0, abs_s = 0.0;
  uint32_t lo_x = 0xffffffffu, hi_x = 0u, lo_q = 0xffffffffu, hi_q = 0u;
#pragma unroll 4
  for (int64_t i = c0 + lane; i < c1; i += 64) {
    const float v = x[i];
    const float sq = v * v;
    const uint32_t ux = __builtin_bit_cast(uint32_t, v) & 0x7fffffffu;
    const uint32_t uq = __builtin_bit_cast(uint32_t, sq) & 0x7fffffffu;
    lo_x = min(lo_x, ux - 1u);  // zero wraps to the largest value
    hi_x = max(hi_x, ux);
    lo_q = min(lo_q, uq - 1u);
    hi_q = max(hi_q, uq);
    s += (double)v;
    q += (double)sq;
    abs_s += fabs((double)v);
  }
  GeChunkSum cs;
  cs.s = s;
  cs.q = q;
  cs.abs_s = abs_s;
  cs.abs_q = q;
  cs.gx = lo_x == 0xffffffffu ? 1000 : ulp_exp_bits(lo_x + 1u);
  cs.gq = lo_q == 0xffffffffu ? 1000 : ulp_exp_bits(lo_q + 1u);
  cs.bad = (hi_x >= 0x7f800000u) | (hi_q >= 0x7f800000u);
  cs.hx = hi_x == 0u ? -1000 : ulp_exp_bits(hi_x);
  reduce_sum_info(cs);
  if (lane == 0) a.csum[c] = cs;
}

// One wave per read: the exactness decision, then the chunks' carries (an
// exclusive scan of their sums, exact when the read is), or the serial pass
// that writes the prefix sums.  Also clears the read's detector mismatch
// marker.
__global__ __launch_bounds__(256) void abea_ge_read_scan_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (r >= a.n_reads) return;
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const int64_t cb = a.chunk_off[r], ce = a.chunk_off[r + 1];
  double *S = a.S + off + r, *Q = a.Q + off + r;
  GeChunkSum t{0.0, 0.0, 0.0, 0.0, 1000, 1000, 0, -1000};
  for (int64_t c = cb + lane; c < ce; c += 64) {
    const GeChunkSum &k = a.csum[c];
    t.abs_s += k.abs_s;
    t.abs_q += k.abs_q;
    t.gx = min(t.gx, k.gx);
    t.gq = min(t.gq, k.gq);
    t.bad |= k.bad;
    t.hx = max(t.hx, k.hx);
  }
  reduce_sum_info(t);
  // the sums of magnitudes are rounded: leave a factor-2 margin
  const bool exact = !t.bad && (t.gx >= 1000 || t.abs_s < ldexp(1.0, 51 + t.gx)) &&
                     (t.gq >= 1000 || t.abs_q < ldexp(1.0, 51 + t.gq));
  // tame: an exact read whose samples are all below 2^60 in magnitude (ulp
  // below 2^37), so that no window sum of squares overflows float, and whose
  // samples and float squares are all zero or multiples of 2^-125 (ulp at
  // least 2^-125), so that every non-zero window sum of either is at least
  // 2^-125 in magnitude
  const bool tame = exact && t.hx <= 36 && t.gx >= -125 && t.gq >= -125;
  if (lane == 0) {
    a.exact[r] = exact ? (tame ? 2 : 1) : 0;
    a.first_bad[r] = INT32_MAX;
    if (!exact) {
      S[0] = 0.0;
      Q[0] = 0.0;
    }
  }
  if (exact) {
    double cs = 0.0, cq = 0.0;
    for (int64_t c0 = cb; c0 < ce; c0 += 64) {
      const int64_t c = c0 + lane;
      const double vs = c < ce ? a.csum[c].s : 0.0, vq = c < ce ? a.csum[c].q : 0.0;
      const double ps = wave_scan_add_f64(vs), pq = wave_scan_add_f64(vq);
      if (c < ce) {
        a.carry[2 * c] = cs + ps - vs;  // exact: every partial sum is representable
        a.carry[2 * c + 1] = cq + pq - vq;
      }
      cs += readlane_f64(ps, 63);
      cq += readlane_f64(pq, 63);
    }
  } else if (lane == 0) {
    // the C source's serial loop, in index order
    const float *x = a.raw + off;
    double s = 0.0, q = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const float sq = x[i] * x[i];
      s = s + (double)x[i];
      q = q + (double)sq;
      S[i + 1] = s;
      Q[i + 1] = q;
    }
  }
}

// --------------------------------------------------------------------------
// t-statistics, the C source's float/double types

// Division by d = 3 or 6 (3 times a power of two) for the statistics of tame
// reads (below): a product by y = RN(1/d) and one fused residual correction,
// Markstein's final step.  In double, RN(1/3) = (1/3)(1 - 2^-54), so x*y
// rounds to within an ulp of x/d, and the corrected quotient is the correctly
// rounded one when it is a normal number (div3k).  Float quotients take the
// same step in float (div3f): equal to IEEE float division over every finite
// float for d = 3, and for d = 6 over every float of magnitude at least
// 2^-125 (smaller ones, whose quotients are subnormal, can round the other
// way), in both cases but for the sign of a zero quotient, which no statistic
// sees (it enters only squares and |m2 - m1|).  A tame read's window sums
// and sums of squares are zero or at least 2^-125 in magnitude; the variance
// may be FLT_MIN, so its quotient by 6 goes through double (div3k, then
// rounded to float: the double quotient is exact enough that the second
// rounding gives the float quotient).  tests/native/div3_check.c compares
// each route with IEEE division.
constexpr double kRcp3 = 1.0 / 3.0, kRcp6 = 1.0 / 6.0;
constexpr float kRcp3f = 1.0f / 3.0f, kRcp6f = 1.0f / 6.0f;

__device__ __forceinline__ double div3k(double x, double d, double y) {
  const double q0 = x * y;
  return __builtin_fma(__builtin_fma(-q0, d, x), y, q0);
}

__device__ __forceinline__ float div3f(float x, float d, float y) {
  const float q0 = x * y;
  return __builtin_fmaf(__builtin_fmaf(-q0, d, x), y, q0);
}

// Correctly rounded square root of a double in [2^-767, 2^1023): the
// sequence LLVM lowers sqrt(double) to on AMDGPU (a hardware reciprocal
// square root estimate, then Goldschmidt and two Newton corrections) without
// the rescaling of smaller inputs and the pass-through of 0 and inf, which
// these inputs never take.
__device__ __forceinline__ double sqrt_mid(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Correctly rounded a / b for a in {0} u [2^-149, 2^128) and b in [2^-65,
// 2^65]: the sequence LLVM lowers a double division to on AMDGPU (hardware
// reciprocal estimate, two Newton steps, quotient and one fused correction)
// without its v_div_scale, v_div_fmas and v_div_fixup steps.  For these
// operands v_div_scale returns its input unchanged (no operand or quotient is
// near the ends of the double range), so v_div_fmas is a plain fma, and
// v_div_fixup changes no finite, non-zero-divisor result; a = 0 gives +0, as
// IEEE division does.
__device__ __forceinline__ double div_mid(double a, double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
  const double r2 = __builtin_fma(r1, __builtin_fma(-b, r1, 1.0), r1);
  const double m = a * r2;
  return __builtin_fma(__builtin_fma(-b, m, a), r2, m);
}

// The statistic from the window before the sample (s1, q1: sums of the
// samples and of their float squares, double) and the window from it (s2, q2,
// rounded to float as the C source does).  TAME: the inputs of a tame read,
// whose divisions by w take the routes above; otherwise IEEE division.  The
// final quotient's numerator |m2 - m1| is a float, and its divisor the
// square root of a float of at least FLT_MIN / 6, so it is in div_mid's
// range.
template <bool TAME>
__device__ __forceinline__ float tstat_tail(double s1, double q1, float s2, float q2, int w) {
  const float wf = (float)w;
  const double wd = (double)w;
  const double yd = w == kWShort ? kRcp3 : kRcp6;
  const float yf = w == kWShort ? kRcp3f : kRcp6f;
  auto d64 = [&](double x) { return TAME ? div3k(x, wd, yd) : x / wd; };
  // window sums and sums of squares (zero or at least 2^-125 in a tame read)
  auto d32 = [&](float x) { return TAME ? div3f(x, wf, yf) : x / wf; };
  // the variance, at least FLT_MIN: the float route only for w = 3
  auto d32v = [&](float x) {
    return TAME ? (w == kWShort ? div3f(x, wf, yf) : (float)div3k((double)x, wd, yd)) : x / wf;
  };
  const float m1 = (float)d64(s1);
  const float m2 = d32(s2);
  double cv = d64(q1);
  cv -= (double)(m1 * m1);
  cv += (double)d32(q2);
  cv -= (double)(m2 * m2);
  const float var = fmaxf((float)cv, FLT_MIN);
  const float dm = m2 - m1;
  const double v = (double)d32v(var);
  return (float)(TAME ? div_mid(fabs((double)dm), sqrt_mid(v)) : fabs((double)dm) / sqrt(v));
}

// The statistic from the prefix sums (a read whose sums are not exact).
__device__ __forceinline__ float tstat_at(const double *S, const double *Q, int64_t n, int w,
                                          int64_t i) {
  if (n < 2 * (int64_t)w || i < w || i > n - w) return 0.0f;
  double s1 = S[i], q1 = Q[i];
  if (i > w) {
    s1 -= S[i - w];
    q1 -= Q[i - w];
  }
  return tstat_tail<false>(s1, q1, (float)(S[i + w] - S[i]), (float)(Q[i + w] - Q[i]), w);
}

// One read's signal as the detector sees it.
struct GeRead {
  const float *x;        // the samples
  const double *S, *Q;   // prefix sums (reads whose sums are not exact)
  int32_t n;
  bool exact;            // window sums of the samples are exact in any order
  bool tame;             // exact, samples below 2^60, samples and squares on a 2^-125 grid
  int64_t avail;         // samples readable from x on: this read's and the later reads'
};

// --------------------------------------------------------------------------
// the peak detector

__device__ __forceinline__ GeState fresh_state() {
  GeState s;
  s.d[0] = GeDet{-1, FLT_MAX, 0, 0};
  s.d[1] = GeDet{-1, FLT_MAX, 0, 0};
  return s;
}

// One sample of both detectors (short first); returns how many boundaries
// it emitted (0-2) in p0, p1.  The C source's branches become selects: a
// detector not masked at i either has no peak (pos == -1: a lower x is the
// new minimum, else a rise of more than the peak height starts a peak at
// i) or has one (a higher x moves it to i; then the short detector masks
// the long one past its peak once the peak clears the threshold, a fall of
// more than the peak height below a peak above the threshold validates it,
// and a valid peak more than w/2 samples back is emitted, the detector
// starting afresh from x).  NaN statistics compare false, as in C.
__device__ __forceinline__ int detect_step(GeState &st, float xs, float xl, int32_t i,
                                           int32_t &p0, int32_t &p1) {
  int ne = 0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    GeDet &d = st.d[k];
    const float x = k == 0 ? xs : xl;
    const float thr = k == 0 ? kThrShort : kThrLong;
    const int32_t w = k == 0 ? kWShort : kWLong;
    const bool act = d.masked_to < i;
    const bool np = d.pos == -1;
    const bool lower = x < d.val;
    const bool rise = !lower && x - d.val > kPeakHeight;
    const bool higher = x > d.val;
    const bool take = act && (np ? (lower || rise) : higher);
    const bool at_i = act && (np ? rise : higher);
    const float val = take ? x : d.val;
    const int32_t pos = at_i ? i : d.pos;
    const bool peak = act && !np;
    if (k == 0) {
      const bool mask = peak && val > thr;
      GeDet &l = st.d[1];
      l.masked_to = mask ? pos + w : l.masked_to;
      l.pos = mask ? -1 : l.pos;
      l.val = mask ? FLT_MAX : l.val;
      l.valid = mask ? 0 : l.valid;
    }
    const int32_t valid = (peak && val - x > kPeakHeight && val > thr) ? 1 : d.valid;
    const bool emit = peak && valid && i - pos > w / 2;
    if (emit) {
      if (ne == 0)
        p0 = pos;
      else
        p1 = pos;
      ++ne;
    }
    d.pos = emit ? -1 : pos;
    d.val = emit ? x : val;
    d.valid = emit ? 0 : valid;
  }
  return ne;
}

// States before sample b are the same for every later sample when they agree
// field by field, masks taken no lower than b - 1 (a mask below b no longer
// masks anything).
__device__ __forceinline__ bool same_state(GeState a, GeState b, int32_t at) {
  bool eq = true;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    eq = eq && a.d[k].pos == b.d[k].pos && a.d[k].valid == b.d[k].valid &&
         __builtin_bit_cast(uint32_t, a.d[k].val) == __builtin_bit_cast(uint32_t, b.d[k].val) &&
         max(a.d[k].masked_to, at - 1) == max(b.d[k].masked_to, at - 1);
  }
  return eq;
}

// Sample j of the read; with CHECK, samples outside the read count as 0
// (where the statistics are 0 anyway).
template <bool CHECK>
__device__ __forceinline__ float ge_x(const GeRead &rd, int32_t j) {
  if (!CHECK) return rd.x[j];
  const float v = rd.x[min(max(j, 0), rd.n - 1)];
  return (uint32_t)j < (uint32_t)rd.n ? v : 0.0f;
}

// Where a lane's boundaries go: its chunk's list pk, entry by entry
// (DirectSink), or through a buffer of its own in LDS that goes out kStage
// entries (one 128-byte line) at a time (StagedSink).  Stored one by one,
// the boundaries of a CU's lanes touch as many lines as there are lanes, and
// the caches write each line back partly filled, many times over.
struct DirectSink {
  int32_t *pk;
  __device__ __forceinline__ void put(int32_t k, int32_t, int32_t p) { pk[k] = p; }
  __device__ __forceinline__ void settle(int32_t, int32_t) {}
  __device__ __forceinline__ void finish(int32_t) {}
};

constexpr int kStage = 32;              // entries per write-out (a power of two)
constexpr int kStageLane = kStage + 1;  // LDS words per lane (odd: no bank conflicts)

// Entries below the multiple of kStage at or under the count are in pk; the
// rest, at most kStage + 1 while a step adds its (at most two) entries, are
// in the lane's buffer from its start.
struct StagedSink {
  int32_t *pk;  // the chunk's list
  int32_t *lb;  // the lane's kStageLane words of LDS
  // entry k of a step that began with `before` entries
  __device__ __forceinline__ void put(int32_t k, int32_t before, int32_t p) {
    lb[k - (before & ~(kStage - 1))] = p;
  }
  __device__ __forceinline__ void settle(int32_t before, int32_t cnt) {
    const int32_t done = before & ~(kStage - 1);
    if ((cnt & ~(kStage - 1)) != done) {
#pragma unroll
      for (int j = 0; j < kStage; ++j) pk[done + j] = lb[j];
      lb[0] = lb[kStage];
    }
  }
  __device__ __forceinline__ void finish(int32_t cnt) {
    const int32_t done = cnt & ~(kStage - 1);
    for (int32_t j = 0; j < cnt - done; ++j) pk[done + j] = lb[j];
  }
};

// One sample of the detector; boundaries at samples >= keep_from go to the
// sink.
template <class Sink>
__device__ __forceinline__ void ge_step(GeState &st, float t1, float t2, int32_t i,
                                        int32_t keep_from, Sink &pk, int32_t &cnt,
                                        GeState *entry) {
  int32_t p0 = 0, p1 = 0;
  const int ne = detect_step(st, t1, t2, i, p0, p1);
  if (i >= keep_from && ne > 0) {
    const int32_t before = cnt;
    pk.put(cnt++, before, p0);
    if (ne > 1) pk.put(cnt++, before, p1);
    pk.settle(before, cnt);
  }
  if (entry && i + 1 == keep_from) *entry = st;
}

// v where c holds, else 0, computed where it stands.  The empty asm
// statements hide the mask's value and pin the result in place: otherwise
// the compiler moves each statistic's computation under a branch of its
// own (the select, or the detector step that uses it), which serialises
// the statistics of an iteration.
__device__ __forceinline__ float keep_if(float v, bool c) {
  uint32_t m = c ? 0xffffffffu : 0u;
  asm volatile("" : "+v"(m));
  float r = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & m);
  asm volatile("" : "+v"(r));
  return r;
}

// An exact read: the statistics from window sums that slide along with the
// samples.  B3 and B6 are the sums (and sums of float squares) of the windows
// from the sample, w = 3 and 6; each step adds the sample that enters them
// and takes off the one that leaves, exactly (every partial sum of an exact
// read is exact).  The window before sample i is the one from i - w, kept in
// a short history (h3, h6).  Six samples per iteration: their twelve
// statistics first, as independent straight-line code that the compiler
// interleaves (each is a long chain of dependent double operations), then
// the six detector steps.  CHECK: an iteration may load past the read's end
// (within 18 samples of it).  TAME: the read is tame (tstat_tail).
template <bool CHECK, bool TAME, class Sink>
__device__ void run_exact(GeState &st, const GeRead &rd, int32_t from, int32_t to,
                          int32_t keep_from, Sink &pk, int32_t &cnt, GeState *entry) {
  const int32_t n = rd.n;
  auto win = [&](int32_t i, int w, double &s, double &q) {
    s = 0.0;
    q = 0.0;
    for (int j = 0; j < w; ++j) {
      const float v = ge_x<true>(rd, i + j);
      s += (double)v;
      q += (double)(v * v);
    }
  };
  // h3[j] = B3 at from + j - 3, h6[j] = B6 at from + j - 6
  double h3s[kWShort], h3q[kWShort], h6s[kWLong], h6q[kWLong];
  double b3s, b3q, b6s, b6q;
#pragma unroll
  for (int j = 0; j < kWShort; ++j) win(from + j - kWShort, kWShort, h3s[j], h3q[j]);
#pragma unroll
  for (int j = 0; j < kWLong; ++j) win(from + j - kWLong, kWLong, h6s[j], h6q[j]);
  win(from, kWShort, b3s, b3q);
  win(from, kWLong, b6s, b6q);
  // xb: samples i0 .. i0 + 11 of the iteration; the next iteration's six
  // new ones load while this one computes
  float xb[2 * kWLong];
#pragma unroll
  for (int j = 0; j < 2 * kWLong; ++j) xb[j] = ge_x<CHECK>(rd, from + j);
  for (int32_t i0 = from; i0 < to; i0 += kWLong) {
    float nx[kWLong];
#pragma unroll
    for (int j = 0; j < kWLong; ++j) nx[j] = ge_x<CHECK>(rd, i0 + 2 * kWLong + j);
    float t1[kWLong], t2[kWLong];
#pragma unroll
    for (int k = 0; k < kWLong; ++k) {
      const int32_t i = i0 + k;
      const int k3 = k % kWShort;
      const float u1 = tstat_tail<TAME>(h3s[k3], h3q[k3], (float)b3s, (float)b3q, kWShort);
      const float u2 = tstat_tail<TAME>(h6s[k], h6q[k], (float)b6s, (float)b6q, kWLong);
      t1[k] = keep_if(u1, n >= 2 * kWShort && i >= kWShort && i <= n - kWShort);
      t2[k] = keep_if(u2, n >= 2 * kWLong && i >= kWLong && i <= n - kWLong);
      h3s[k3] = b3s;
      h3q[k3] = b3q;
      h6s[k] = b6s;
      h6q[k] = b6q;
      const float x0 = xb[k], x3 = xb[k + kWShort], x6 = xb[k + kWLong];
      const double d0 = (double)x0, e0 = (double)(x0 * x0);
      b3s = (b3s + (double)x3) - d0;
      b3q = (b3q + (double)(x3 * x3)) - e0;
      b6s = (b6s + (double)x6) - d0;
      b6q = (b6q + (double)(x6 * x6)) - e0;
    }
#pragma unroll
    for (int k = 0; k < kWLong; ++k)
      if (i0 + k < to) ge_step(st, t1[k], t2[k], i0 + k, keep_from, pk, cnt, entry);
#pragma unroll
    for (int j = 0; j < kWLong; ++j) {
      xb[j] = xb[j + kWLong];
      xb[j + kWLong] = nx[j];
    }
  }
}

// Run samples [from, to) of a read from state st; boundaries emitted at
// samples >= keep_from go to pk (count in *np), the state before keep_from
// to *entry.  The statistics come with the samples: from sliding window sums
// (exact reads) or from the prefix sums.
template <class Sink>
__device__ void run_detector(GeState &st, const GeRead &rd, int32_t from, int32_t to,
                             int32_t keep_from, Sink &pk, int32_t *np, GeState *entry) {
  int32_t cnt = 0;
  if (entry && from == keep_from) *entry = st;
  if (rd.exact) {
    // An iteration loads samples up to i0 + 17 < to + 3 * 6.  Past the
    // read's end those are the next read's samples: they enter only window
    // sums whose statistics keep_if zeroes (samples above n - w), so only
    // the end of the last read needs the checked loads, and a wave whose
    // lanes reach a read's end runs one variant, not two.
    const bool check = (int64_t)to + 3 * kWLong > rd.avail;
    if (rd.tame) {
      if (check)
        run_exact<true, true, Sink>(st, rd, from, to, keep_from, pk, cnt, entry);
      else
        run_exact<false, true, Sink>(st, rd, from, to, keep_from, pk, cnt, entry);
    } else {
      if (check)
        run_exact<true, false, Sink>(st, rd, from, to, keep_from, pk, cnt, entry);
      else
        run_exact<false, false, Sink>(st, rd, from, to, keep_from, pk, cnt, entry);
    }
  } else {
    for (int32_t i = from; i < to; ++i) {
      const float t1 = tstat_at(rd.S, rd.Q, rd.n, kWShort, i);
      const float t2 = tstat_at(rd.S, rd.Q, rd.n, kWLong, i);
      ge_step(st, t1, t2, i, keep_from, pk, cnt, entry);
    }
  }
  pk.finish(cnt);
  *np = cnt;
}

// One lane per detector chunk: a fresh run from a.overlap samples before the
// chunk.  A chunk of D samples holds at most 2 * (ceil(D / 3) + 1) <= D
// boundaries (D >= 8): a detector that emits is back to no peak and needs 3
// more samples to emit again.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kGeDetBlocksPerCU))) void abea_ge_detect_kernel(GeArgs a) {
  __shared__ int32_t stage[kStageLane * 256];  // each lane's boundary buffer
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_det) return;
  const int r = a.det_read[c];
  const int64_t off = a.raw_off[r];
  const int32_t n = (int32_t)(a.raw_off[r + 1] - off);
  const int32_t c0 = (int32_t)(c - a.det_off[r]) * a.det_chunk;
  const int32_t c1 = min(c0 + a.det_chunk, n);
  const int32_t from = max(c0 - a.overlap, 0);
  GeState st = fresh_state();
  const GeRead rd{a.raw + off, a.S + off + r, a.Q + off + r, n, a.exact[r] != 0,
                  a.exact[r] == 2, a.raw_off[a.n_reads] - off};
  StagedSink sink{a.peaks + c * a.det_chunk, stage + kStageLane * threadIdx.x};
  run_detector(st, rd, from, c1, c0, sink, a.n_peaks + c, a.entry + c);
  a.exit[c] = st;
}

// One thread per detector chunk: is its entry its predecessor's exit?
__global__ __launch_bounds__(256) void abea_ge_check_kernel(GeArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_det) return;
  const int r = a.det_read[c];
  const int32_t lc = (int32_t)(c - a.det_off[r]);
  if (lc == 0) return;  // the first chunk starts from the true initial state
  if (!same_state(a.entry[c], a.exit[c - 1], lc * a.det_chunk)) atomicMin(&a.first_bad[r], lc);
}

// One lane per read.  The chunks before the first mismatch ran from the true
// state (induction from the first chunk); from there on, each chunk whose
// entry is not the true state is re-run from it.
__global__ __launch_bounds__(256) void abea_ge_fix_kernel(GeArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const int32_t fb = a.first_bad[r];
  int32_t reruns = 0;
  if (fb != INT32_MAX) {
    const int64_t off = a.raw_off[r];
    const int32_t n = (int32_t)(a.raw_off[r + 1] - off);
    const int64_t cb = a.det_off[r];
    GeState st = a.exit[cb + fb - 1];
    const GeRead rd{a.raw + off, a.S + off + r, a.Q + off + r, n, a.exact[r] != 0,
                    a.exact[r] == 2, a.raw_off[a.n_reads] - off};
    for (int64_t c = cb + fb; c < a.det_off[r + 1]; ++c) {
      const int32_t c0 = (int32_t)(c - cb) * a.det_chunk;
      if (same_state(a.entry[c], st, c0)) {
        st = a.exit[c];
      } else {
        ++reruns;
        DirectSink sink{a.peaks + c * a.det_chunk};
        run_detector(st, rd, c0, min(c0 + a.det_chunk, n), c0, sink, a.n_peaks + c, nullptr);
      }
    }
  }
  a.reruns[r] = reruns;
}

// One wave per read (4 per block): where each detector chunk's events go.
// Chunk c's boundaries are the read's boundaries ev_base[c] .. ev_base[c] +
// n_peaks[c] - 1 (an exclusive scan of the counts), and boundary g ends event
// g; the event that ends at its first boundary starts at prev_last[c], the
// last boundary in the chunks before it (an exclusive maximum scan:
// boundaries increase along the read; 0 where there is none).  Also the
// event count: m boundaries give max(m, 1) events, scrappie's create_events
// count (the last boundary ends no event; the last event runs to the end).
__global__ __launch_bounds__(256) void abea_ge_scan_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (r >= a.n_reads) return;
  const int64_t n = a.raw_off[r + 1] - a.raw_off[r];
  const int64_t cb = a.det_off[r], ce = a.det_off[r + 1];
  int32_t base = 0, last = 0;
  for (int64_t c0 = cb; c0 < ce; c0 += 64) {
    const int64_t c = c0 + lane;
    const int32_t cnt = c < ce ? a.n_peaks[c] : 0;
    const int32_t lst = cnt > 0 ? a.peaks[(int64_t)c * a.det_chunk + cnt - 1] : 0;
    const int incl = wave_scan_add(cnt);
    int32_t mx = lst;  // inclusive maximum scan
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t v = __shfl_up(mx, o, 64);
      if (lane >= o) mx = max(mx, v);
    }
    const int32_t before = __shfl_up(mx, 1, 64);
    if (c < ce) {
      a.ev_base[c] = base + incl - cnt;
      a.prev_last[c] = max(last, lane > 0 ? before : 0);
    }
    base += __shfl(incl, 63, 64);
    last = max(last, __shfl(mx, 63, 64));
  }
  if (lane == 0) a.n_events[r] = n > 0 ? max(base, 1) : 0;
}

// --------------------------------------------------------------------------
// events

__device__ __forceinline__ abea_event_t make_event(int64_t start, int64_t end, const double *S,
                                                   const double *Q) {
  abea_event_t e{};
  e.start = (uint64_t)start;
  e.length = (float)((uint64_t)end - (uint64_t)start);  // wraps if end < start, as size_t does
  e.mean = (float)(S[end] - S[start]) / e.length;
  const float dsq = (float)(Q[end] - Q[start]);
  const float var = dsq / e.length - e.mean * e.mean;
  e.stdv = sqrtf(fmaxf(var, 0.0f));
  return e;
}

// Exact sums of samples [0, p) of an exact read (sum of samples, of their
// float squares): the carry of p's chunk plus the samples before p in it.
__device__ __forceinline__ void exact_prefix(const GeArgs &a, const float *x, int64_t cbase,
                                             int64_t p, double &s, double &q) {
  s = 0.0;
  q = 0.0;
  if (p <= 0) return;
  const int64_t lc = (p - 1) / kGeChunk, k0 = lc * kGeChunk;
  s = a.carry[2 * (cbase + lc)];
  q = a.carry[2 * (cbase + lc) + 1];
  for (int64_t k = k0; k < p; ++k) {
    const float v = x[k];
    s += (double)v;
    q += (double)(v * v);
  }
}

// An event from the exact sums of its samples (read exact) or from the
// prefix sums (serial reads); the same values either way.
__device__ __forceinline__ abea_event_t make_event_exact(const GeArgs &a, const float *x,
                                                         int64_t cbase, int64_t start,
                                                         int64_t end) {
  double ds, dq;
  if (start <= end && end - start <= 2 * kGeChunk) {
    // four partial sums (exact in any order).  The first 16 samples, which
    // hold most events whole, load together; the rest four per step.
    // Missing samples count as +0.0, which changes no sum.
    double s[4] = {0.0, 0.0, 0.0, 0.0}, q[4] = {0.0, 0.0, 0.0, 0.0};
    float v0[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v0[u] = start + u < end ? x[start + u] : 0.0f;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      s[u & 3] += (double)v0[u];
      q[u & 3] += (double)(v0[u] * v0[u]);
    }
    for (int64_t k = start + 16; k < end; k += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float v = k + u < end ? x[k + u] : 0.0f;
        s[u] += (double)v;
        q[u] += (double)(v * v);
      }
    }
    ds = (s[0] + s[1]) + (s[2] + s[3]);
    dq = (q[0] + q[1]) + (q[2] + q[3]);
  } else {
    double s0, q0, s1, q1;
    exact_prefix(a, x, cbase, start, s0, q0);
    exact_prefix(a, x, cbase, end, s1, q1);
    ds = s1 - s0;
    dq = q1 - q0;
  }
  abea_event_t e{};
  e.start = (uint64_t)start;
  e.length = (float)((uint64_t)end - (uint64_t)start);
  e.mean = (float)ds / e.length;
  const float dsq = (float)dq;
  const float var = dsq / e.length - e.mean * e.mean;
  e.stdv = sqrtf(fmaxf(var, 0.0f));
  return e;
}

// One wave per detector chunk (4 per block): the events its boundaries end.
// Event j runs from boundary j - 1 (0 for the first) to boundary j, except
// the read's last event, which runs from its second-last boundary to the end
// of the signal: its last boundary ends no event (scrappie's create_events,
// whose event count is the boundary count).  A read with no boundary gets
// one event, written by its last chunk.
__global__ __launch_bounds__(256) void abea_ge_events_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= a.n_det) return;
  const int r = a.det_read[c];
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const int32_t cnt = a.n_peaks[c];
  const int64_t lc = a.det_off[r + 1] - 1;  // the read's last chunk
  const int64_t nb = (int64_t)a.ev_base[lc] + a.n_peaks[lc];  // the read's boundaries
  const int32_t m = (nb == 0 && c == lc) ? 1 : cnt;
  const int32_t *pk = a.peaks + c * a.det_chunk;
  const int64_t jb = a.ev_base[c];
  const int32_t pl = a.prev_last[c];
  abea_event_t *ev = a.events + off + r;
  const bool exact = a.exact[r] != 0;
  const int64_t cbase = a.chunk_off[r];
  const double *S = a.S + off + r, *Q = a.Q + off + r;
  const float *x = a.raw + off;
  for (int32_t k = lane; k < m; k += 64) {
    const int64_t start = k == 0 ? pl : pk[k - 1];
    const int64_t end = jb + k + 1 < nb ? pk[k] : n;
    ev[jb + k] = exact ? make_event_exact(a, x, cbase, start, end) : make_event(start, end, S, Q);
  }
}

}  // namespace

hipError_t launch_getevents(const GeArgs &a, hipStream_t s) {
  if (a.n_reads <= 0) return hipSuccess;
  const unsigned rb = (unsigned)((a.n_reads + 3) / 4);     // a wave per read
  const unsigned cb = (unsigned)((a.n_chunks + 3) / 4);    // a wave per prefix chunk
  const unsigned db = (unsigned)((a.n_det + 255) / 256);   // a lane per detector chunk
  if (a.n_chunks > 0) hipLaunchKernelGGL(abea_ge_chunk_sum_kernel, dim3(cb), dim3(256), 0, s, a);
  hipLaunchKernelGGL(abea_ge_read_scan_kernel, dim3(rb), dim3(256), 0, s, a);
  if (a.n_det > 0) {
    hipLaunchKernelGGL(abea_ge_detect_kernel, dim3(db), dim3(256), 0, s, a);
    hipLaunchKernelGGL(abea_ge_check_kernel, dim3(db), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(abea_ge_fix_kernel, dim3((unsigned)((a.n_reads + 255) / 256)), dim3(256), 0,
                     s, a);
  hipLaunchKernelGGL(abea_ge_scan_kernel, dim3(rb), dim3(256), 0, s, a);
  if (a.n_det > 0)
    hipLaunchKernelGGL(abea_ge_events_kernel, dim3((unsigned)((a.n_det + 3) / 4)), dim3(256), 0,
                       s, a);
  return hipGetLastError();
}

}  // namespace abea
