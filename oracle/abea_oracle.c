/*
 * abea_oracle.c -- CPU restatement of f5c's ABEA (see abea_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Never linked into the product library.
 *
 * Built with -ffp-contract=off so that every a*b+c below rounds twice, as the
 * x86-64 f5c build does.
 */
#include "abea_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define W 100            /* ALN_BANDWIDTH, GPU/f5c.h:30 */
#define HALF_W 50
#define KMER 6           /* KMER_SIZE, GPU/f5c.h:26 */
#define FROM_D 0         /* GPU/align_cuda.cl:130-132 */
#define FROM_U 1
#define FROM_L 2
#define MAX_GAP_THRESHOLD 50          /* GPU/align_cuda.cl:134 */
#define MIN_AVG_LOG_EMISSION (-5.0)   /* GPU/align_cuda.cl:139 */
#define LOG_INV_SQRT_2PI (-0.918938f) /* GPU/align_cuda.cl:62 */

void orc_transition_constants(int32_t n_events, int32_t n_kmers, double out[4]) {
  /* align_swi.cl:233-250: events_per_kmer and p_stay are double on the CPU. */
#ifdef ORC_MUTATE_FLOAT_PSTAY /* planted mistake: GPU kernel's float p_stay */
  float events_per_kmer = (float)n_events / n_kmers;
  float p_stay = 1 - (1 / (events_per_kmer + 1));
#else
  double events_per_kmer = (double)n_events / n_kmers;
  double p_stay = 1 - (1 / (events_per_kmer + 1));
#endif
  double lp_skip = log(1e-10);
  double lp_stay = log(p_stay);
  double lp_step = log(1.0 - exp(lp_skip) - exp(lp_stay));
  double lp_trim = log(0.01);
  out[0] = lp_skip;
  out[1] = lp_stay;
  out[2] = lp_step;
  out[3] = lp_trim;
}

static inline uint32_t base_code(char b) {
  /* GPU/align_cuda.cl:22-35 */
  switch (b) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return 0;
  }
}

uint32_t orc_kmer_rank(const char *s) {
  /* GPU/align_cuda.cl:39-51: r += rank(str[k-i-1]) << 2i, so s[0] is the
   * most significant base-4 digit. */
  uint32_t r = 0;
#ifdef ORC_MUTATE_RANK_LSB_FIRST /* planted mistake: reversed digit order */
  for (uint32_t i = 0; i < KMER; ++i) r += base_code(s[i]) << (i << 1);
#else
  for (uint32_t i = 0; i < KMER; ++i) r += base_code(s[KMER - i - 1]) << (i << 1);
#endif
  return r;
}

float orc_emission_f32(float x, const orc_model_t *m, float scale, float shift) {
  /* GPU/align_cuda.cl:511-528 with CACHED_LOG: the cached level_log_stdv is
   * used as-is; scaling.var does not enter (gp_stdv = level_stdv * 1). */
  float gp_mean = scale * m->level_mean + shift;
  float gp_stdv = m->level_stdv;
#ifdef ORC_MUTATE_RECOMPUTE_LOG /* planted mistake: log recomputed */
  float gp_log_stdv = logf(m->level_stdv);
#else
  float gp_log_stdv = m->level_log_stdv;
#endif
  float a = (x - gp_mean) / gp_stdv;
  return LOG_INV_SQRT_2PI - gp_log_stdv + (-0.5f * a * a);
}

static inline double emission_f64(float x, const orc_model_t *m, float scale, float shift) {
  double gp_mean = (double)scale * (double)m->level_mean + (double)shift;
  double a = ((double)x - gp_mean) / (double)m->level_stdv;
  return (double)LOG_INV_SQRT_2PI - (double)m->level_log_stdv + (-0.5 * a * a);
}

/* ------------------------------------------------------------------------ */
/* The band fill is written once over a cell type: float for the f5c
 * semantics, double for the error oracle. */

#define CELL float
#define ORC_FILL orc_fill_f32
#define EMIT(x, m, sc, sh) orc_emission_f32((x), (m), (sc), (sh))
#define ROUND(v) ((float)(v))
#include "abea_oracle_fill.inc"
#undef CELL
#undef ORC_FILL
#undef EMIT
#undef ROUND

#define CELL double
#define ORC_FILL orc_fill_f64
#define EMIT(x, m, sc, sh) emission_f64((x), (m), (sc), (sh))
#define ROUND(v) ((double)(v))
#include "abea_oracle_fill.inc"
#undef CELL
#undef ORC_FILL
#undef EMIT
#undef ROUND

int32_t orc_align(orc_pair_t *out, int32_t out_cap, const char *seq, int32_t seq_len,
                  const float *ev_mean, int32_t n_events, const orc_model_t *models,
                  float scale, float shift, int mode, const uint8_t *replay_moves,
                  uint8_t *moves_out, orc_pair_t *path_out, uint8_t *from_out,
                  uint8_t *trace_out, float *endcol_out, orc_info_t *info) {
  if (mode == ORC_MODE_FP64)
    return orc_fill_f64(out, out_cap, seq, seq_len, ev_mean, n_events, models, scale, shift,
                        replay_moves, moves_out, path_out, from_out, trace_out, endcol_out, info);
  return orc_fill_f32(out, out_cap, seq, seq_len, ev_mean, n_events, models, scale, shift,
                      replay_moves, moves_out, path_out, from_out, trace_out, endcol_out, info);
}

void orc_path_gradient(const orc_pair_t *path, const uint8_t *from, int32_t n, const char *seq,
                       const float *ev_mean, const orc_model_t *models, float scale, float shift,
                       double weight, double *g_read, double *g_mu, double *g_sigma,
                       double *g_lambda) {
  /* score = sum over D/U path cells of em(e,k) + terms that do not depend on
   * the scalings or the model.  With a = (x - (scale*mu + shift)) / sigma and
   * em = c - lambda - a^2/2 (lambda = the cached level_log_stdv):
   *   d em/d shift = a/sigma, d em/d scale = a*mu/sigma,
   *   d em/d mu = a*scale/sigma, d em/d sigma = a^2/sigma (lambda held),
   *   d em/d lambda = -1,
   *   d em/d var = 0 (GPU/align_cuda.cl:95: gp_stdv = level_stdv * 1). */
  double gs = 0, gc = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (from[i] == FROM_L) continue;
    int32_t k = path[i].ref_pos, e = path[i].read_pos;
    uint32_t r = orc_kmer_rank(seq + k);
    const orc_model_t *m = &models[r];
    double mu = m->level_mean, sg = m->level_stdv;
    double a = ((double)ev_mean[e] - ((double)scale * mu + (double)shift)) / sg;
    gs += a / sg;
    gc += a * mu / sg;
    if (g_mu) g_mu[r] += weight * (a * (double)scale / sg);
    if (g_sigma) g_sigma[r] += weight * (a * a / sg);
    if (g_lambda) g_lambda[r] -= weight;
  }
  if (g_read) {
    g_read[0] += weight * gs;
    g_read[1] += weight * gc;
    g_read[2] += 0.0;
  }
}

double orc_path_score(const orc_pair_t *path, const uint8_t *from, int32_t n, const char *seq,
                      const float *ev_mean, int32_t n_events, const orc_model_t *models,
                      float scale, float shift) {
  /* FP64 score of a fixed path: boundary cell + transitions + emissions of
   * D/U cells + trimming of the events after the end cell.  Equals the DP's
   * best score when the path is the DP's own (GPU/align_cuda.cl:532-542,
   * 722-746). */
  if (n <= 0) return -INFINITY;
  int32_t n_kmers = (int32_t)strlen(seq) - KMER + 1;
  double lp[4];
  orc_transition_constants(n_events, n_kmers, lp);
  const double lp_skip = lp[0], lp_stay = lp[1], lp_step = lp[2], lp_trim = lp[3];
  int32_t k0 = path[0].ref_pos, e0 = path[0].read_pos;
  int32_t pk = from[0] == FROM_U ? k0 : k0 - 1;
  int32_t pe = from[0] == FROM_L ? e0 : e0 - 1;
  double s;
  if (pk == -1 && pe == -1) s = 0.0;
  else if (pk == -1) s = lp_trim * (pe + 1);
  else return -INFINITY;
  for (int32_t i = 0; i < n; ++i) {
    int32_t k = path[i].ref_pos, e = path[i].read_pos;
    if (from[i] == FROM_L) {
      s += lp_skip;
    } else {
      s += from[i] == FROM_D ? lp_step : lp_stay;
      s += emission_f64(ev_mean[e], &models[orc_kmer_rank(seq + k)], scale, shift);
    }
  }
  return s + (n_events - path[n - 1].read_pos) * lp_trim;
}

typedef struct { int64_t key; int32_t idx; } sort_item;

static int cmp_desc(const void *a, const void *b) {
  const sort_item *x = (const sort_item *)a, *y = (const sort_item *)b;
  if (x->key != y->key) return x->key > y->key ? -1 : 1;
  return x->idx - y->idx;
}

int orc_align_batch(int32_t n_reads, const char *seqs, const int64_t *seq_ptr,
                    const int32_t *seq_len, const float *ev_means, const int64_t *ev_ptr,
                    const int32_t *n_events, const orc_model_t *models, const float *scale,
                    const float *shift, int mode, const orc_batch_out_t *o, int n_threads) {
  /* The reference's CPU path is OpenMP/pthreads over reads
   * (GPU/f5c.h:57-59); longest reads first keeps the threads balanced. */
  sort_item *order = (sort_item *)malloc(sizeof(sort_item) * (size_t)(n_reads > 0 ? n_reads : 1));
  for (int32_t i = 0; i < n_reads; ++i) {
    order[i].key = (int64_t)n_events[i] + seq_len[i];
    order[i].idx = i;
  }
  qsort(order, (size_t)n_reads, sizeof(sort_item), cmp_desc);
  int used = 1;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
#endif
  for (int32_t j = 0; j < n_reads; ++j) {
    int32_t i = order[j].idx;
    orc_info_t info;
    int32_t cap = n_events[i] + (seq_len[i] > 5 ? seq_len[i] - 5 : 0);
    int64_t off = o->pair_ptr ? o->pair_ptr[i] : 0;
    orc_pair_t *pr = o->pairs ? o->pairs + off : NULL;
    orc_pair_t *path = o->path ? o->path + off : NULL;
    uint8_t *from = o->from ? o->from + off : NULL;
    int32_t np = orc_align(pr, pr ? cap : 0, seqs + seq_ptr[i], seq_len[i], ev_means + ev_ptr[i],
                           n_events[i], models, scale[i], shift[i], mode, NULL, NULL, path, from,
                           NULL, NULL, &info);
    if (o->n_pairs) o->n_pairs[i] = np;
    if (o->path_len) o->path_len[i] = info.path_len;
    if (o->score) o->score[i] = info.score;
    if (o->avg) o->avg[i] = info.avg_log_emission;
    if (o->fills) o->fills[i] = info.fills;
  }
#ifdef _OPENMP
  used = n_threads;
#endif
  free(order);
  return used;
}

void orc_fills_from_moves(int32_t n_reads, const int32_t *n_events, const int32_t *n_kmers,
                          const uint32_t *moves, const int64_t *mv_off, int64_t *fills) {
  /* The fill range of each band, align_swi.cl:420-432, replayed from the
   * moves: band b >= 2 moves right (K + 1) or down (E + 1) from band b - 1;
   * band 1's lower-left cell is (E, K) = (50, -51), align_swi.cl:306-322. */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int32_t i = 0; i < n_reads; ++i) {
    const int32_t ne = n_events[i], nk = n_kmers[i];
    int64_t f = 0;
    if (ne > 0 && nk > 0) {
      const int64_t nb = (int64_t)ne + nk + 2;
      const uint32_t *mv = moves + mv_off[i];
      int64_t E = HALF_W, K = -1 - HALF_W;
      for (int64_t b = 2; b < nb; ++b) {
        if ((mv[b >> 5] >> (b & 31)) & 1u) ++K; else ++E;
        int64_t lo = -K, emin = E - (ne - 1);
        if (emin > lo) lo = emin;
        if (lo < 0) lo = 0;
        int64_t hi = nk - K;
        if (E + 1 < hi) hi = E + 1;
        if (hi > W) hi = W;
        if (hi > lo) f += hi - lo;
      }
    }
    fills[i] = f;
  }
}
