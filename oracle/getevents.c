/*
 * getevents.c -- CPU restatement of f5c's getevents(): event detection on a
 * read's raw signal (picoamperes), the step before
 * estimate_scalings_using_mom() and align().
 *
 * TEST INFRASTRUCTURE ONLY (the checker for libabea's event detection
 * kernels); the product library never links or calls it.
 *
 * The reference declares getevents(nsample, rawptr) (GPU/f5cmisc.h:44) and
 * its types, which it marks as adapted from scrappie (event_t and
 * event_table, GPU/f5c.h:112-130).  The body lives in f5c (fork imsuneth/f5c,
 * F5C_VERSION "0.4", GPU/f5c.h:23), which the reference does not vendor and
 * does not pin.  f5c takes it, by way of nanopolish, from scrappie's event
 * detection (nanoporetech/scrappie, event_detection.c) with scrappie's DNA
 * defaults; the trimming call made before detection in that code returns a
 * trimmed copy that is not used, so detection runs on the whole signal.
 * This file restates the published algorithm:
 *
 *  1. Prefix sums S[i] = x[0] + ... + x[i-1] and Q[i] of the squares, in
 *     double and index order; each square is a float product.
 *  2. Two t-statistics, window lengths w = 3 and w = 6.  For i in [w, n-w]
 *     the windows x[i-w..i) and x[i..i+w) give means m1, m2 and population
 *     variances; t[i] = |m2 - m1| / sqrt(max(v1 + v2, FLT_MIN) / w).  t is 0
 *     elsewhere, and everywhere when n < 2w.
 *  3. The short/long peak detector (thresholds 1.4 and 9.0, peak height
 *     0.2).  Each detector follows a running minimum until its statistic
 *     rises more than the peak height above it, then a running maximum (the
 *     peak).  A peak above the threshold from which the statistic has fallen
 *     more than the peak height is valid; a valid peak more than w/2 samples
 *     old is emitted as an event boundary and the detector restarts from
 *     the current value.  While the short detector holds a peak above its
 *     threshold, the long detector is masked to w_short samples past that
 *     peak and restarted.
 *  4. Events from the boundaries b_0 < ... < b_{m-1} as scrappie's
 *     create_events counts them: nevent = 1 + #{i >= 1 : 0 < b_i < n}, so m
 *     boundaries give m events: [0, b_0), [b_0, b_1), ..., [b_{m-3}, b_{m-2})
 *     and the last one [b_{m-2}, n).  The last boundary ends no event.
 *     Start, length, mean and stdv from the prefix sums, in float.
 *
 * The mixed float/double arithmetic follows the C source's types, noted at
 * each step.  With fewer than two boundaries that code reads before its
 * boundary array (b_{-1}); here the one event then covers the whole signal,
 * as b_{-1} = 0 would give.  The reference's fixtures hold no raw signal, so
 * parity against f5c is unpinned; tests/test_getevents_oracle.py pins this
 * restatement by closed forms.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "abea_oracle.h"

/* scrappie's event_detection_defaults (DNA) */
#define W_SHORT 3
#define W_LONG 6
#define THR_SHORT 1.4f
#define THR_LONG 9.0f
#define PEAK_HEIGHT 0.2f

void orc_signal_prefix(const float *x, int64_t n, double *S, double *Q) {
  S[0] = 0.0;
  Q[0] = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const float sq = x[i] * x[i]; /* float * float, then widened */
#ifdef ORC_MUTATE_GE_FLOAT_PREFIX
    S[i + 1] = (float)(S[i] + (double)x[i]);
#else
    S[i + 1] = S[i] + (double)x[i];
#endif
    Q[i + 1] = Q[i] + (double)sq;
  }
}

void orc_tstat(const double *S, const double *Q, int64_t n, int32_t w, float *t) {
  for (int64_t i = 0; i < n; ++i) t[i] = 0.0f;
  if (n < 2 * (int64_t)w || w < 2) return;
  const float wf = (float)w;
  for (int64_t i = w; i <= n - w; ++i) {
    double s1 = S[i], q1 = Q[i];
    if (i > w) {
      s1 -= S[i - w];
      q1 -= Q[i - w];
    }
    const float s2 = (float)(S[i + w] - S[i]);
    const float q2 = (float)(Q[i + w] - Q[i]);
    const float m1 = (float)(s1 / (double)wf); /* double / float -> double */
    const float m2 = s2 / wf;                  /* float / float */
    /* q1/w - m1*m1 + q2/w - m2*m2, left to right: double, float product,
     * float quotient, float product */
    double cv = q1 / (double)wf;
    cv -= (double)(m1 * m1);
    cv += (double)(q2 / wf);
    cv -= (double)(m2 * m2);
    const float var = fmaxf((float)cv, FLT_MIN);
    const float dm = m2 - m1;
    t[i] = (float)(fabs((double)dm) / sqrt((double)(var / wf)));
  }
}

typedef struct {
  const float *sig;
  float thr;
  int64_t w;
  int64_t masked_to; /* skips samples i <= masked_to */
  int64_t pos;       /* -1: no peak yet */
  float val;         /* running minimum (no peak) or the peak's value */
  int valid;
} orc_detector_t;

static void detector_init(orc_detector_t *d, const float *sig, float thr, int64_t w) {
  d->sig = sig;
  d->thr = thr;
  d->w = w;
  d->masked_to = 0;
  d->pos = -1;
  d->val = FLT_MAX;
  d->valid = 0;
}

int64_t orc_detect_peaks(const float *t_short, const float *t_long, int64_t n, int64_t *peaks) {
  orc_detector_t ds, dl;
  detector_init(&ds, t_short, THR_SHORT, W_SHORT);
  detector_init(&dl, t_long, THR_LONG, W_LONG);
  orc_detector_t *det[2] = {&ds, &dl};
  int64_t np = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 2; ++k) {
      orc_detector_t *d = det[k];
      if (d->masked_to >= i) continue;
      const float x = d->sig[i];
      if (d->pos == -1) {
        if (x < d->val) {
          d->val = x;
        } else if (x - d->val > PEAK_HEIGHT) {
          d->val = x;
          d->pos = i;
        }
      } else {
        if (x > d->val) {
          d->val = x;
          d->pos = i;
        }
#ifdef ORC_MUTATE_GE_NO_MASK
        if (0) {
#else
        if (d == &ds && d->val > d->thr) {
#endif
          dl.masked_to = d->pos + d->w;
          dl.pos = -1;
          dl.val = FLT_MAX;
          dl.valid = 0;
        }
        if (d->val - x > PEAK_HEIGHT && d->val > d->thr) d->valid = 1;
#ifdef ORC_MUTATE_GE_EMIT_EARLY
        if (d->valid) {
#else
        if (d->valid && i - d->pos > d->w / 2) {
#endif
          peaks[np++] = d->pos;
          d->pos = -1;
          d->val = x;
          d->valid = 0;
        }
      }
    }
  }
  return np;
}

static orc_event_t make_event(int64_t start, int64_t end, const double *S, const double *Q) {
  orc_event_t e;
  e.start = (uint64_t)start;
  /* an unsigned difference, converted to float (wraps if end < start) */
  e.length = (float)((uint64_t)end - (uint64_t)start);
  e.mean = (float)(S[end] - S[start]) / e.length;
  const float dsq = (float)(Q[end] - Q[start]);
  const float var = dsq / e.length - e.mean * e.mean;
  e.stdv = sqrtf(fmaxf(var, 0.0f));
  return e;
}

int64_t orc_getevents(const float *raw, int64_t n, orc_event_t *out, int64_t cap) {
  if (n <= 0) return 0;
  double *S = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *Q = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  float *t1 = (float *)malloc(sizeof(float) * (size_t)n);
  float *t2 = (float *)malloc(sizeof(float) * (size_t)n);
  int64_t *peaks = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  int64_t m = 0; /* 0: allocation failed (a signal has at least one event) */
  if (S && Q && t1 && t2 && peaks) {
    orc_signal_prefix(raw, n, S, Q);
    orc_tstat(S, Q, n, W_SHORT, t1);
    orc_tstat(S, Q, n, W_LONG, t2);
    const int64_t np = orc_detect_peaks(t1, t2, n, peaks);
#ifdef ORC_MUTATE_GE_KEEP_LAST_BOUNDARY /* planted mistake: every boundary ends an event */
    m = np + 1;
#else
    m = np > 1 ? np : 1;
#endif
    if (m <= cap) {
      int64_t prev = 0;
      for (int64_t j = 0; j + 1 < m; ++j) {
        out[j] = make_event(prev, peaks[j], S, Q);
        prev = peaks[j];
      }
      out[m - 1] = make_event(prev, n, S, Q);
    } else {
      m = -m; /* capacity too small: minus the event count */
    }
  }
  free(S);
  free(Q);
  free(t1);
  free(t2);
  free(peaks);
  return m;
}
