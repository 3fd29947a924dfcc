/*
 * postalign.c -- CPU restatement of the f5c steps either side of align():
 * estimate_scalings_using_mom() before it, postalign() and
 * recalibrate_model() after it.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for libabea's postalign kernels);
 * the product library never links or calls it.
 *
 * The reference declares these functions (GPU/f5cmisc.h:47, :51-52, :53-54)
 * and their types (index_pair_t and event_alignment_t, GPU/f5c.h:184-207;
 * the read status flags, GPU/f5c.h:49-51), but their bodies live in f5c
 * (fork imsuneth/f5c, F5C_VERSION "0.4", GPU/f5c.h:23), which the reference
 * does not vendor and does not pin to a commit.  f5c ports them from
 * nanopolish (SquiggleRead::load_from_raw, get_eventalignment_for_1d_basecalls,
 * recalibrate_model, estimate_scalings_using_mom); this file restates that
 * published algorithm:
 *
 *  - estimate_scalings_using_mom: shift = mean(event means) - mean(k-mer
 *    level means); scale = mean((x - shift)^2) / mean(level^2), all sums in
 *    double, in index order; var and log_var stay 0.  Pinned bit for bit by
 *    the scalings f5c dumped for the 143 reads of
 *    FPGA/code/bins/dump_small/0 (they are the values f5c passed to align).
 *  - postalign: each event on the path is assigned to the k-mer of its first
 *    pair; base_to_event_map[k] = {first, last} assigned event or {-1, -1};
 *    events_per_base = (max event - min event) / n_kmers; one record per
 *    assigned event in k-mer order, hmm_state 'M' when the k-mer rank differs
 *    from the previous record's, else 'E'.
 *  - recalibrate_model: weighted least squares of event mean on level mean
 *    (weights 1/level_stdv^2) over the 'M' records, in record order, solved
 *    by Cramer's rule; with scale_var, var = sqrt(mean of squared weighted
 *    residuals) and log_var = log(var); needs at least 200 'M' records.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "abea_oracle.h"

#define KMER 6
#define MIN_M_RECORDS 200

static uint32_t rank6(const char *s) { return orc_kmer_rank(s); }

orc_scalings_t orc_estimate_scalings_mom(const char *sequence, int32_t sequence_len,
                                         const orc_model_t *model, const float *ev_mean,
                                         int32_t n_events) {
  const int32_t n_kmers = sequence_len - KMER + 1;
  double ev_sum = 0.0;
  for (int32_t i = 0; i < n_events; ++i) ev_sum += ev_mean[i];
  double km_sum = 0.0, km_sq = 0.0;
  for (int32_t i = 0; i < n_kmers; ++i) {
    double l = model[rank6(sequence + i)].level_mean;
    km_sum += l;
    km_sq += l * l;
  }
  double shift = ev_sum / n_events - km_sum / n_kmers;
  double ev_sq = 0.0;
  for (int32_t i = 0; i < n_events; ++i) {
#ifdef ORC_MUTATE_MOM_FLOAT_RESIDUAL /* planted mistake: residual rounded to float */
    float d = (float)(ev_mean[i] - shift);
    ev_sq += (double)d * d;
#else
    ev_sq += (ev_mean[i] - shift) * (ev_mean[i] - shift);
#endif
  }
  double scale = (ev_sq / n_events) / (km_sq / n_kmers);
  orc_scalings_t out;
  out.shift = (float)shift;
  out.scale = (float)scale;
  out.var = 0.0f;
  out.log_var = 0.0f;
  return out;
}

int32_t orc_postalign(orc_event_alignment_t *alignment, orc_index_pair_t *base_to_event_map,
                      double *events_per_base, const char *sequence, int32_t n_kmers,
                      const orc_pair_t *pairs, int32_t n_pairs) {
  for (int32_t k = 0; k < n_kmers; ++k) base_to_event_map[k].start = base_to_event_map[k].stop = -1;
  int32_t max_event = 0, min_event = INT32_MAX, prev_event = -1;
  for (int32_t i = 0; i < n_pairs; ++i) {
    const int32_t k = pairs[i].ref_pos, e = pairs[i].read_pos;
    orc_index_pair_t *m = &base_to_event_map[k];
#ifdef ORC_MUTATE_POSTALIGN_EVERY_PAIR /* planted mistake: repeated events also assigned */
    if (1) {
#else
    if (e != prev_event) {
#endif
      if (m->start == -1) m->start = e;
      m->stop = e;
    }
    if (e > max_event) max_event = e;
    if (e < min_event) min_event = e;
    prev_event = e;
  }
  *events_per_base = (double)(max_event - min_event) / n_kmers;

  int32_t n = 0, prev_rank = -1;
  for (int32_t k = 0; k < n_kmers; ++k) {
    const orc_index_pair_t r = base_to_event_map[k];
    if (r.start == -1) continue;
    const int32_t rank = (int32_t)rank6(sequence + k);
    for (int32_t e = r.start; e <= r.stop; ++e) {
      orc_event_alignment_t *a = &alignment[n++];
      memset(a, 0, sizeof *a);
      memcpy(a->ref_kmer, sequence + k, KMER);
      memcpy(a->model_kmer, sequence + k, KMER);
      a->ref_position = k;
      a->read_idx = -1;
      a->event_idx = e;
      a->rc = 0;
#ifdef ORC_MUTATE_POSTALIGN_STATE_BY_POSITION /* planted mistake: M on a new k-mer index */
      a->hmm_state = (e == r.start) ? 'M' : 'E';
#else
      a->hmm_state = prev_rank != rank ? 'M' : 'E';
#endif
      prev_rank = rank;
    }
  }
  return n;
}

int orc_recalibrate_model(const orc_model_t *model, const float *ev_mean, orc_scalings_t *sc,
                          const orc_event_alignment_t *alignment, int32_t n_alignment,
                          int scale_var) {
  int32_t n_m = 0;
  for (int32_t i = 0; i < n_alignment; ++i) n_m += alignment[i].hmm_state == 'M';
  if (n_m < MIN_M_RECORDS) return 0;
  double a00 = 0, a01 = 0, a11 = 0, b0 = 0, b1 = 0;
  for (int32_t i = 0; i < n_alignment; ++i) {
    const orc_event_alignment_t *a = &alignment[i];
#ifndef ORC_MUTATE_RECAL_ALL_RECORDS /* planted mistake: 'E' records counted too */
    if (a->hmm_state != 'M') continue;
#endif
    const orc_model_t *m = &model[rank6(a->ref_kmer)];
    const double x = ev_mean[a->event_idx];
    const double mu = m->level_mean, sd = m->level_stdv;
    const double w = 1.0 / (sd * sd);
    a00 += w;
    a01 += mu * w;
    a11 += mu * mu * w;
    b0 += x * w;
    b1 += mu * x * w;
  }
  const double a10 = a01;
  const double det = a00 * a11 - a01 * a10;
  const double shift = -(a01 * b1 - a11 * b0) / det;
  const double scale = (a00 * b1 - a10 * b0) / det;
  double var = 1.0;
  if (scale_var) {
    var = 0.0;
    for (int32_t i = 0; i < n_alignment; ++i) {
      const orc_event_alignment_t *a = &alignment[i];
      if (a->hmm_state != 'M') continue;
      const orc_model_t *m = &model[rank6(a->ref_kmer)];
      const double x = ev_mean[a->event_idx];
      const double mu = m->level_mean, sd = m->level_stdv;
      const double y = x - shift - scale * mu;
      var += y * y / (sd * sd);
    }
    var /= n_m;
    var = sqrt(var);
  }
  sc->shift = (float)shift;
  sc->scale = (float)scale;
  sc->var = (float)var;
  sc->log_var = (float)log(var);
  return 1;
}
