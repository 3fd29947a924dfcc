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
//   tstat    one thread per sample (window sums from the chunk's samples in
//            LDS), the C source's float/double types.
//   detect   one lane per 512-sample chunk, started fresh a few samples
//            early (speculation): the chunk's entry state, its boundaries in
//            emission order and its exit state.
//   check / fix   a chunk is right when its entry state is its
//            predecessor's exit state (the detector is a deterministic state
//            machine and the first chunk starts from the true initial
//            state); from a read's first chunk where that fails, one lane
//            walks on, re-running each chunk whose entry is not the true
//            state.  The result is the serial one.
//   compact / events   boundaries gathered per read, one thread per event.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "abea_internal.h"
#include "abea_wave.h"

namespace abea {

namespace {

constexpr float kThrShort = 1.4f, kThrLong = 9.0f, kPeakHeight = 0.2f;
constexpr int kWShort = 3, kWLong = 6;

// The t-statistics of read r start at toff(r): 128-byte (cache line)
// aligned, after the previous read's end (align32(b) - align32(a) + 32 >
// b - a).  The detector's block loads read up to 15 floats past a read's
// end: into the next read's statistics, or into the 64 floats allocated past
// the last read's.
__device__ __forceinline__ int64_t toff(const GeArgs &a, int r) {
  return ((a.raw_off[r] + 31) & ~(int64_t)31) + 32 * (int64_t)r;
}

// --------------------------------------------------------------------------
// prefix sums

// ulp exponent of a float (-149 for subnormals); 1000 for zero (no
// constraint), and 2000 for inf/NaN (never exact)
__device__ __forceinline__ int ulp_exp(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f) & 0x7fffffffu;
  if (u == 0) return 1000;
  const int e = (int)(u >> 23);
  if (e == 255) return 2000;
  return e == 0 ? -149 : e - 150;
}

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
}

// One wave per prefix chunk (4 per block): the chunk's sums and exactness
// data, read 64 consecutive samples per step.
__global__ __launch_bounds__(256) void abea_ge_chunk_sum_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= a.n_chunks) return;
  const int r = a.chunk_read[c];
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const int64_t c0 = (c - a.chunk_off[r]) * kGeChunk, c1 = min(c0 + kGeChunk, n);
  const float *x = a.raw + off;
  GeChunkSum cs{0.0, 0.0, 0.0, 0.0, 1000, 1000, 0, 0};
#pragma unroll 4
  for (int64_t i = c0 + lane; i < c1; i += 64) {
    const float v = x[i];
    const float sq = v * v;
    const int ex = ulp_exp(v), eq = ulp_exp(sq);
    cs.bad |= (ex == 2000) | (eq == 2000);
    cs.gx = min(cs.gx, ex);
    cs.gq = min(cs.gq, eq);
    cs.s += (double)v;
    cs.q += (double)sq;
    cs.abs_s += fabs((double)v);
    cs.abs_q += (double)sq;
  }
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
  GeChunkSum t{0.0, 0.0, 0.0, 0.0, 1000, 1000, 0, 0};
  for (int64_t c = cb + lane; c < ce; c += 64) {
    const GeChunkSum &k = a.csum[c];
    t.abs_s += k.abs_s;
    t.abs_q += k.abs_q;
    t.gx = min(t.gx, k.gx);
    t.gq = min(t.gq, k.gq);
    t.bad |= k.bad;
  }
  reduce_sum_info(t);
  // the sums of magnitudes are rounded: leave a factor-2 margin
  const bool exact = !t.bad && (t.gx >= 1000 || t.abs_s < ldexp(1.0, 51 + t.gx)) &&
                     (t.gq >= 1000 || t.abs_q < ldexp(1.0, 51 + t.gq));
  if (lane == 0) {
    a.exact[r] = exact ? 1 : 0;
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
// t-statistics, one thread per sample; the C source's float/double types
__device__ __forceinline__ float tstat_at(const double *S, const double *Q, int64_t n, int w,
                                          int64_t i) {
  if (n < 2 * (int64_t)w || i < w || i > n - w) return 0.0f;
  const float wf = (float)w;
  double s1 = S[i], q1 = Q[i];
  if (i > w) {
    s1 -= S[i - w];
    q1 -= Q[i - w];
  }
  const float s2 = (float)(S[i + w] - S[i]);
  const float q2 = (float)(Q[i + w] - Q[i]);
  const float m1 = (float)(s1 / (double)wf);
  const float m2 = s2 / wf;
  double cv = q1 / (double)wf;
  cv -= (double)(m1 * m1);
  cv += (double)(q2 / wf);
  cv -= (double)(m2 * m2);
  const float var = fmaxf((float)cv, FLT_MIN);
  const float dm = m2 - m1;
  return (float)(fabs((double)dm) / sqrt((double)(var / wf)));
}

// The statistic from exact window sums (a read whose sums are exact): the
// same values as tstat_at's prefix-sum differences.  xs[k] is sample k - 6
// of the chunk's halo-extended block.
__device__ __forceinline__ float tstat_win(const float *xs, int64_t n, int w, int64_t i, int k) {
  if (n < 2 * (int64_t)w || i < w || i > n - w) return 0.0f;
  const float wf = (float)w;
  double s1 = 0.0, q1 = 0.0, s2d = 0.0, q2d = 0.0;
  for (int j = 0; j < w; ++j) {
    const float a = xs[k - w + j], b = xs[k + j];
    s1 += (double)a;
    q1 += (double)(a * a);
    s2d += (double)b;
    q2d += (double)(b * b);
  }
  const float s2 = (float)s2d, q2 = (float)q2d;
  const float m1 = (float)(s1 / (double)wf);
  const float m2 = s2 / wf;
  double cv = q1 / (double)wf;
  cv -= (double)(m1 * m1);
  cv += (double)(q2 / wf);
  cv -= (double)(m2 * m2);
  const float var = fmaxf((float)cv, FLT_MIN);
  const float dm = m2 - m1;
  return (float)(fabs((double)dm) / sqrt((double)(var / wf)));
}

// One block per prefix chunk.
__global__ __launch_bounds__(256) void abea_ge_tstat_kernel(GeArgs a) {
  __shared__ float xs[kGeChunk + 2 * kWLong];
  const int64_t c = blockIdx.x;
  const int r = a.chunk_read[c];
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const int64_t c0 = (c - a.chunk_off[r]) * kGeChunk, c1 = min(c0 + kGeChunk, n);
  const int64_t to = toff(a, r);
  if (a.exact[r]) {
    const float *x = a.raw + off;
    for (int k = threadIdx.x; k < kGeChunk + 2 * kWLong; k += blockDim.x) {
      const int64_t i = c0 - kWLong + k;
      xs[k] = (i >= 0 && i < n) ? x[i] : 0.0f;
    }
    __syncthreads();
    for (int64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
      const int k = (int)(i - c0) + kWLong;
      a.t1[to + i] = tstat_win(xs, n, kWShort, i, k);
      a.t2[to + i] = tstat_win(xs, n, kWLong, i, k);
    }
  } else {
    const double *S = a.S + off + r, *Q = a.Q + off + r;
    for (int64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
      a.t1[to + i] = tstat_at(S, Q, n, kWShort, i);
      a.t2[to + i] = tstat_at(S, Q, n, kWLong, i);
    }
  }
}

// --------------------------------------------------------------------------
// the peak detector

__device__ __forceinline__ GeState fresh_state() {
  GeState s;
  s.d[0] = GeDet{-1, FLT_MAX, 0, 0};
  s.d[1] = GeDet{-1, FLT_MAX, 0, 0};
  return s;
}

// One sample of both detectors (short first); returns how many boundaries
// it emitted (0-2) in p0, p1.
__device__ __forceinline__ int detect_step(GeState &st, float xs, float xl, int32_t i,
                                           int32_t &p0, int32_t &p1) {
  int ne = 0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    GeDet &d = st.d[k];
    if (d.masked_to >= i) continue;
    const float x = k == 0 ? xs : xl;
    const float thr = k == 0 ? kThrShort : kThrLong;
    const int32_t w = k == 0 ? kWShort : kWLong;
    if (d.pos == -1) {
      if (x < d.val) {
        d.val = x;
      } else if (x - d.val > kPeakHeight) {
        d.val = x;
        d.pos = i;
      }
    } else {
      if (x > d.val) {
        d.val = x;
        d.pos = i;
      }
      if (k == 0 && d.val > thr) {
        GeDet &l = st.d[1];
        l.masked_to = d.pos + w;
        l.pos = -1;
        l.val = FLT_MAX;
        l.valid = 0;
      }
      if (d.val - x > kPeakHeight && d.val > thr) d.valid = 1;
      if (d.valid && i - d.pos > w / 2) {
        if (ne == 0)
          p0 = d.pos;
        else
          p1 = d.pos;
        ++ne;
        d.pos = -1;
        d.val = x;
        d.valid = 0;
      }
    }
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

// Run samples [from, to) of a read from state st; boundaries emitted at
// samples >= keep_from go to pk (count in *np).  from is a multiple of 16 and
// t1, t2 are cache-line aligned: the statistics come in blocks of 16 samples
// (float4 loads), the next block loading while this one runs.
__device__ void run_detector(GeState &st, const float *t1, const float *t2, int32_t from,
                             int32_t to, int32_t keep_from, int32_t *pk, int32_t *np,
                             GeState *entry) {
  int32_t cnt = 0;
  if (entry && from == keep_from) *entry = st;
  const float4 *q1 = reinterpret_cast<const float4 *>(t1), *q2 = reinterpret_cast<const float4 *>(t2);
  float4 n1[4], n2[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    n1[k] = q1[from / 4 + k];
    n2[k] = q2[from / 4 + k];
  }
  for (int32_t i0 = from; i0 < to; i0 += 16) {
    float b1[16], b2[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      b1[4 * k] = n1[k].x, b1[4 * k + 1] = n1[k].y, b1[4 * k + 2] = n1[k].z, b1[4 * k + 3] = n1[k].w;
      b2[4 * k] = n2[k].x, b2[4 * k + 1] = n2[k].y, b2[4 * k + 2] = n2[k].z, b2[4 * k + 3] = n2[k].w;
    }
    if (i0 + 16 < to) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        n1[k] = q1[(i0 + 16) / 4 + k];
        n2[k] = q2[(i0 + 16) / 4 + k];
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int32_t i = i0 + k;
      if (i < to) {
        int32_t p0 = 0, p1 = 0;
        const int ne = detect_step(st, b1[k], b2[k], i, p0, p1);
        if (i >= keep_from) {
          if (ne > 0) pk[cnt++] = p0;
          if (ne > 1) pk[cnt++] = p1;
        }
        if (entry && i + 1 == keep_from) *entry = st;
      }
    }
  }
  *np = cnt;
}

// One lane per detector chunk: a fresh run from a.overlap samples before the
// chunk.  At most 2 * (ceil(512 / 3) + 1) = 344 boundaries: a detector that
// emits is back to no peak and needs 3 more samples to emit again.
__global__ __launch_bounds__(256) void abea_ge_detect_kernel(GeArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_det) return;
  const int r = a.det_read[c];
  const int64_t off = a.raw_off[r];
  const int32_t n = (int32_t)(a.raw_off[r + 1] - off);
  const int32_t c0 = (int32_t)(c - a.det_off[r]) * kGeDetChunk;
  const int32_t c1 = min(c0 + kGeDetChunk, n);
  const int32_t from = max(c0 - a.overlap, 0) & ~15;  // block-aligned (a longer lead-in)
  GeState st = fresh_state();
  const int64_t to = toff(a, r);
  run_detector(st, a.t1 + to, a.t2 + to, from, c1, c0, a.peaks + c * kGeDetChunk,
               a.n_peaks + c, a.entry + c);
  a.exit[c] = st;
}

// One thread per detector chunk: is its entry its predecessor's exit?
__global__ __launch_bounds__(256) void abea_ge_check_kernel(GeArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_det) return;
  const int r = a.det_read[c];
  const int32_t lc = (int32_t)(c - a.det_off[r]);
  if (lc == 0) return;  // the first chunk starts from the true initial state
  if (!same_state(a.entry[c], a.exit[c - 1], lc * kGeDetChunk)) atomicMin(&a.first_bad[r], lc);
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
    for (int64_t c = cb + fb; c < a.det_off[r + 1]; ++c) {
      const int32_t c0 = (int32_t)(c - cb) * kGeDetChunk;
      if (same_state(a.entry[c], st, c0)) {
        st = a.exit[c];
      } else {
        ++reruns;
        run_detector(st, a.t1 + toff(a, r), a.t2 + toff(a, r), c0, min(c0 + kGeDetChunk, n), c0,
                     a.peaks + c * kGeDetChunk, a.n_peaks + c, nullptr);
      }
    }
  }
  a.reruns[r] = reruns;
}

// One wave per read (4 per block): the boundaries gathered in emission order
// (into the read's t1 space, free by now) and the event count.
__global__ __launch_bounds__(256) void abea_ge_compact_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (r >= a.n_reads) return;
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  int32_t *bd = a.bounds + off;
  const int64_t cb = a.det_off[r], ce = a.det_off[r + 1];
  int64_t base = 0;
  for (int64_t c0 = cb; c0 < ce; c0 += 64) {
    const int64_t c = c0 + lane;
    const int32_t cnt = c < ce ? a.n_peaks[c] : 0;
    const int incl = wave_scan_add(cnt);
    const int64_t j = base + incl - cnt;
    for (int32_t k = 0; k < cnt; ++k) bd[j + k] = a.peaks[c * kGeDetChunk + k];
    base += __shfl(incl, 63, 64);
  }
  if (lane == 0) a.n_events[r] = n > 0 ? base + 1 : 0;
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
    ds = 0.0;
    dq = 0.0;
    for (int64_t k = start; k < end; ++k) {
      const float v = x[k];
      ds += (double)v;
      dq += (double)(v * v);
    }
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

// One block per prefix chunk: events j in [2048 lc, 2048 (lc + 1)) of the
// read (the last chunk takes any rest), one thread each.  Event j runs from
// boundary j - 1 (0 for the first) to boundary j (n for the last).
__global__ __launch_bounds__(256) void abea_ge_events_kernel(GeArgs a) {
  const int64_t c = blockIdx.x;
  const int r = a.chunk_read[c];
  const int64_t cbase = a.chunk_off[r];
  const int64_t lc = c - cbase, nch = a.chunk_off[r + 1] - cbase;
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const int64_t m = a.n_events[r];
  const int64_t j0 = lc * kGeChunk, j1 = lc == nch - 1 ? m : min(j0 + kGeChunk, m);
  const int32_t *bd = a.bounds + off;
  abea_event_t *ev = a.events + off + r;
  const bool exact = a.exact[r] != 0;
  const double *S = a.S + off + r, *Q = a.Q + off + r;
  const float *x = a.raw + off;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const int64_t start = j == 0 ? 0 : bd[j - 1];
    const int64_t end = j == m - 1 ? n : bd[j];
    ev[j] = exact ? make_event_exact(a, x, cbase, start, end) : make_event(start, end, S, Q);
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
  if (a.n_chunks > 0)
    hipLaunchKernelGGL(abea_ge_tstat_kernel, dim3((unsigned)a.n_chunks), dim3(256), 0, s, a);
  if (a.n_det > 0) {
    hipLaunchKernelGGL(abea_ge_detect_kernel, dim3(db), dim3(256), 0, s, a);
    hipLaunchKernelGGL(abea_ge_check_kernel, dim3(db), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(abea_ge_fix_kernel, dim3((unsigned)((a.n_reads + 255) / 256)), dim3(256), 0,
                     s, a);
  hipLaunchKernelGGL(abea_ge_compact_kernel, dim3(rb), dim3(256), 0, s, a);
  if (a.n_chunks > 0)
    hipLaunchKernelGGL(abea_ge_events_kernel, dim3((unsigned)a.n_chunks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace abea
