// Event detection for gfx950: f5c's getevents() (GPU/f5cmisc.h:44), the
// step before estimate_scalings_using_mom() and align().  The reference
// declares it and its types (event_t / event_table, adapted from scrappie,
// GPU/f5c.h:112-130); the body is f5c's port of scrappie's detector:
// prefix sums of the signal and its squares (double), two windowed
// t-statistics (w = 3 and 6), the short/long peak detector (thresholds 1.4
// and 9.0, peak height 0.2), and events between consecutive boundaries.
//
//   prefix   one wave per read.  The sums are exact in any order when every
//            term is a multiple of 2^g and the total stays below 2^(52+g)
//            (true for pA-scale signals); then a DPP scan per 64 samples,
//            else a serial pass in index order.
//   tstat    one thread per sample, the C source's float/double types.
//   detect   one lane per 2048-sample chunk, started fresh 256 samples early
//            (speculation): the chunk's entry state, its boundaries in
//            emission order and its exit state.
//   verify   one lane per read walks its chunks in order: a chunk whose entry
//            state equals the true state there is accepted as is (the
//            detector is a deterministic state machine), otherwise it is
//            re-run from the true state.  The result is the serial one.
//   events   one wave per read: boundaries -> event records.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "abea_internal.h"
#include "abea_wave.h"

namespace abea {

namespace {

constexpr float kThrShort = 1.4f, kThrLong = 9.0f, kPeakHeight = 0.2f;
constexpr int kWShort = 3, kWLong = 6;

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

__global__ __launch_bounds__(256) void abea_ge_prefix_kernel(GeArgs a) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (r >= a.n_reads) return;
  const int64_t off = a.raw_off[r];
  const int64_t n = a.raw_off[r + 1] - off;
  const float *x = a.raw + off;
  double *S = a.S + off + r, *Q = a.Q + off + r;
  // exactness: every term (and square) a multiple of 2^g, total < 2^(52+g)
  int gx = 1000, gq = 1000, bad = 0;
  double ax = 0.0, aq = 0.0;
  for (int64_t i = lane; i < n; i += 64) {
    const float v = x[i];
    const float sq = v * v;
    const int ex = ulp_exp(v), eq = ulp_exp(sq);
    bad |= (ex == 2000) | (eq == 2000);
    gx = min(gx, ex);
    gq = min(gq, eq);
    ax += fabs((double)v);
    aq += (double)sq;
  }
  gx = wave_min_i32(gx);
  gq = wave_min_i32(gq);
  bad = wave_max_i32(bad);
  ax = wave_sum_any_f64(ax);
  aq = wave_sum_any_f64(aq);
  // the sums of magnitudes are rounded: leave a factor-2 margin
  const bool exact = !bad && (gx >= 1000 || ax < ldexp(1.0, 51 + gx)) &&
                     (gq >= 1000 || aq < ldexp(1.0, 51 + gq));
  if (lane == 0) {
    S[0] = 0.0;
    Q[0] = 0.0;
  }
  if (exact) {
    double cs = 0.0, cq = 0.0;  // running totals (exact)
    for (int64_t b = 0; b < n; b += 64) {
      const int64_t i = b + lane;
      const float v = i < n ? x[i] : 0.0f;
      const float sq = v * v;
      const double ps = wave_scan_add_f64((double)v), pq = wave_scan_add_f64((double)sq);
      if (i < n) {
        S[i + 1] = cs + ps;
        Q[i + 1] = cq + pq;
      }
      cs += readlane_f64(ps, 63);
      cq += readlane_f64(pq, 63);
    }
  } else if (lane == 0) {
    // the C source's serial loop, in index order
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

__global__ __launch_bounds__(256) void abea_ge_tstat_kernel(GeArgs a) {
  for (int r = blockIdx.x; r < a.n_reads; r += gridDim.x) {
    const int64_t off = a.raw_off[r];
    const int64_t n = a.raw_off[r + 1] - off;
    const double *S = a.S + off + r, *Q = a.Q + off + r;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      a.t1[off + i] = tstat_at(S, Q, n, kWShort, i);
      a.t2[off + i] = tstat_at(S, Q, n, kWLong, i);
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
// samples >= keep_from go to pk (count in *np).
__device__ void run_detector(GeState &st, const float *t1, const float *t2, int32_t from,
                             int32_t to, int32_t keep_from, int32_t *pk, int32_t *np,
                             GeState *entry) {
  int32_t cnt = 0;
  if (entry && from == keep_from) *entry = st;
  for (int32_t i = from; i < to; ++i) {
    int32_t p0 = 0, p1 = 0;
    const int ne = detect_step(st, t1[i], t2[i], i, p0, p1);
    if (i >= keep_from) {
      if (ne > 0) pk[cnt++] = p0;
      if (ne > 1) pk[cnt++] = p1;
    }
    if (entry && i + 1 == keep_from) *entry = st;
  }
  *np = cnt;
}

__global__ __launch_bounds__(256) void abea_ge_detect_kernel(GeArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_chunks) return;
  const int r = a.chunk_read[c];
  const int64_t off = a.raw_off[r];
  const int32_t n = (int32_t)(a.raw_off[r + 1] - off);
  const int32_t c0 = (int32_t)(c - a.chunk_off[r]) * kGeChunk;
  const int32_t c1 = min(c0 + kGeChunk, n);
  const int32_t from = max(c0 - a.overlap, 0);
  GeState st = fresh_state();
  run_detector(st, a.t1 + off, a.t2 + off, from, c1, c0, a.peaks + c * kGeChunk, a.n_peaks + c,
               a.entry + c);
  a.exit[c] = st;
}

__global__ __launch_bounds__(256) void abea_ge_verify_kernel(GeArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const int64_t off = a.raw_off[r];
  const int32_t n = (int32_t)(a.raw_off[r + 1] - off);
  GeState st = fresh_state();
  int32_t reruns = 0;
  for (int64_t c = a.chunk_off[r]; c < a.chunk_off[r + 1]; ++c) {
    const int32_t c0 = (int32_t)(c - a.chunk_off[r]) * kGeChunk;
    if (same_state(a.entry[c], st, c0)) {
      st = a.exit[c];
    } else {
      ++reruns;
      run_detector(st, a.t1 + off, a.t2 + off, c0, min(c0 + kGeChunk, n), c0,
                   a.peaks + c * kGeChunk, a.n_peaks + c, nullptr);
    }
  }
  a.reruns[r] = reruns;
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

__global__ __launch_bounds__(64) void abea_ge_events_kernel(GeArgs a) {
  const int lane = threadIdx.x;
  for (int r = blockIdx.x; r < a.n_reads; r += gridDim.x) {
    const int64_t off = a.raw_off[r];
    const int64_t n = a.raw_off[r + 1] - off;
    const double *S = a.S + off + r, *Q = a.Q + off + r;
    abea_event_t *ev = a.events + off + r;
    const int64_t cb = a.chunk_off[r], ce = a.chunk_off[r + 1];
    int64_t base = 0;       // events before this round of chunks
    int32_t last_pk = 0;    // the last boundary of the chunks before it (0: none)
    for (int64_t c0 = cb; c0 < ce; c0 += 64) {
      const int64_t c = c0 + lane;
      const int32_t cnt = c < ce ? a.n_peaks[c] : 0;
      const int32_t *pk = a.peaks + c * kGeChunk;
      const int incl = wave_scan_add(cnt);
      // the last boundary before this chunk: the last one of the nearest
      // earlier chunk with any, in this round, else the carried one
      const int own = cnt > 0 ? lane : -1;
      const int below = __shfl_up(wave_scan_max(own), 1, 64);
      int32_t prev = last_pk;
      if (lane > 0 && below >= 0) prev = a.peaks[(c0 + below) * kGeChunk + a.n_peaks[c0 + below] - 1];
      int64_t j = base + incl - cnt;
      for (int32_t k = 0; k < cnt; ++k, ++j) {
        ev[j] = make_event(prev, pk[k], S, Q);
        prev = pk[k];
      }
      const int top = wave_scan_max(own);
      const int lastc = __shfl(top, 63, 64);
      if (lastc >= 0) last_pk = a.peaks[(c0 + lastc) * kGeChunk + a.n_peaks[c0 + lastc] - 1];
      base += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
      if (n > 0) ev[base] = make_event(last_pk, n, S, Q);
      a.n_events[r] = n > 0 ? base + 1 : 0;
    }
  }
}

}  // namespace

hipError_t launch_getevents(const GeArgs &a, hipStream_t s) {
  if (a.n_reads <= 0) return hipSuccess;
  hipLaunchKernelGGL(abea_ge_prefix_kernel, dim3((a.n_reads + 3) / 4), dim3(256), 0, s, a);
  const int grid = a.n_reads < 4096 ? a.n_reads : 4096;
  hipLaunchKernelGGL(abea_ge_tstat_kernel, dim3(grid), dim3(256), 0, s, a);
  if (a.n_chunks > 0) {
    hipLaunchKernelGGL(abea_ge_detect_kernel, dim3((unsigned)((a.n_chunks + 255) / 256)), dim3(256),
                       0, s, a);
  }
  hipLaunchKernelGGL(abea_ge_verify_kernel, dim3((a.n_reads + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(abea_ge_events_kernel, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace abea
