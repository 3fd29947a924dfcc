/*
 * abea_oracle.h -- CPU restatement of f5c's Adaptive Banded Event Alignment.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library (libabea) never links or calls it.
 *
 * Semantics follow the reference's CPU-shaped sources:
 *   fill       FPGA/code/device/align_swi.cl:230-503
 *   traceback  FPGA/code/host/src/align_swi.cpp:1139-1312
 *   emission   GPU/align_cuda.cl:62-110 (cached level_log_stdv, CACHED_LOG
 *              GPU/f5c.h:67)
 *   types      GPU/f5c.h:112-181
 * with the precision choices of upstream f5c align(): double transition
 * constants and p_stay, float band cells, double QC emission sum.
 */
#ifndef ABEA_ORACLE_H
#define ABEA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float level_mean, level_stdv, level_log_stdv; } orc_model_t; /* GPU/f5c.h:133-148 */
typedef struct { int32_t ref_pos, read_pos; } orc_pair_t;                     /* GPU/f5c.h:177-181 */

enum { ORC_MODE_F5C = 0,  /* f5c mixed precision (float cells, double transitions) */
       ORC_MODE_FP64 = 1  /* every quantity in double: the error oracle */ };

typedef struct {
  double score;             /* best end score (float-valued in ORC_MODE_F5C) */
  double avg_log_emission;  /* QC statistic, GPU/align_cuda.cl:868 */
  int32_t path_len;         /* pairs on the traced path before QC */
  int32_t max_gap;
  int32_t spanned;
  int32_t qc_pass;
  int32_t end_event;        /* e* of the best end cell */
  int32_t n_bands;
  int64_t fills;            /* cells actually computed */
} orc_info_t;

/* Per-read transition log-probabilities, align_swi.cl:233-250. out[4] =
 * {lp_skip, lp_stay, lp_step, lp_trim}. */
void orc_transition_constants(int32_t n_events, int32_t n_kmers, double out[4]);

/* 6-mer rank, base-4 with S[k] the most significant digit,
 * GPU/align_cuda.cl:22-51. Non-ACGT maps to 0. */
uint32_t orc_kmer_rank(const char *s);

/* f5c-shaped emission (float), GPU/align_cuda.cl:511-528. */
float orc_emission_f32(float x, const orc_model_t *m, float scale, float shift);

/*
 * Align one read.  Returns the number of pairs written to out (0 when QC
 * fails, or when the read has fewer than 6 bases / no events).
 *
 * out, out_cap     ascending AlignedPairs after QC (may be NULL)
 * replay_moves     if non-NULL, n_bands bytes: 1 = band b moved right,
 *                  0 = down, used instead of Suzuki's rule (band held fixed)
 * moves_out        if non-NULL, n_bands bytes of the moves taken
 * path_out/from_out  pre-QC path in ascending order with from codes
 *                  (0 = D, 1 = U, 2 = L); capacity n_events + n_kmers
 * trace_out        if non-NULL, n_bands*100 trace bytes (reference layout)
 * endcol_out       if non-NULL, n_events floats: band cell (e, n_kmers-1),
 *                  -inf when outside the band
 */
int32_t orc_align(orc_pair_t *out, int32_t out_cap,
                  const char *seq, int32_t seq_len,
                  const float *ev_mean, int32_t n_events,
                  const orc_model_t *models, float scale, float shift,
                  int mode, const uint8_t *replay_moves, uint8_t *moves_out,
                  orc_pair_t *path_out, uint8_t *from_out,
                  uint8_t *trace_out, float *endcol_out, orc_info_t *info);

/*
 * FP64 gradient of the best-path score along a given path (from codes as
 * returned by orc_align): only D and U cells carry an emission.  The partial
 * derivatives of the function the DP computes, which reads level_log_stdv as
 * its own input (CACHED_LOG, GPU/align_cuda.cl:517-518):
 * g_read[0..2] += d/d(shift, scale, var)   (var term is identically zero)
 * g_mu[rank]     += d/d level_mean
 * g_sigma[rank]  += d/d level_stdv with level_log_stdv held
 * g_lambda[rank] += d/d level_log_stdv
 * With level_log_stdv = log(level_stdv) the total derivative in level_stdv
 * is g_sigma + g_lambda / level_stdv.  Any output may be NULL.
 */
void orc_path_gradient(const orc_pair_t *path, const uint8_t *from, int32_t n,
                       const char *seq, const float *ev_mean,
                       const orc_model_t *models, float scale, float shift,
                       double weight, double *g_read, double *g_mu,
                       double *g_sigma, double *g_lambda);

/* FP64 score of a fixed path (from codes as returned by orc_align). */
double orc_path_score(const orc_pair_t *path, const uint8_t *from, int32_t n, const char *seq,
                      const float *ev_mean, int32_t n_events, const orc_model_t *models,
                      float scale, float shift);

/* Batched CPU baseline and parity reference: OpenMP over reads, longest
 * first.  seq_ptr / ev_ptr are CSR offsets (int64) into seqs / ev_means.
 * Read i's slots start at out->pair_ptr[i] (capacity n_events + n_kmers);
 * every output array may be NULL.  mode is ORC_MODE_F5C or ORC_MODE_FP64.
 * Returns the number of threads used. */
typedef struct {
  orc_pair_t *pairs;        /* ascending pairs after QC, at pair_ptr[i] */
  orc_pair_t *path;         /* ascending traced path before QC, at pair_ptr[i] */
  uint8_t *from;            /* its from codes (0 D, 1 U, 2 L), at pair_ptr[i] */
  const int64_t *pair_ptr;  /* needed when pairs, path or from is given */
  int32_t *n_pairs;         /* pairs after QC (0 = failed) */
  int32_t *path_len;        /* path length before QC */
  double *score;            /* best end score (float-valued in ORC_MODE_F5C) */
  double *avg;              /* QC statistic */
  int64_t *fills;           /* cells computed (the reference's fills counter) */
} orc_batch_out_t;

int orc_align_batch(int32_t n_reads, const char *seqs, const int64_t *seq_ptr,
                    const int32_t *seq_len, const float *ev_means,
                    const int64_t *ev_ptr, const int32_t *n_events,
                    const orc_model_t *models, const float *scale,
                    const float *shift, int mode, const orc_batch_out_t *out,
                    int n_threads);

/* Cells computed (fills) of each read from its band moves, packed as the
 * library's move words: read i's band b is bit b % 32 of word
 * mv_off[i] + b / 32 (1 = right).  Reads with no k-mers or events get 0. */
void orc_fills_from_moves(int32_t n_reads, const int32_t *n_events, const int32_t *n_kmers,
                          const uint32_t *moves, const int64_t *mv_off, int64_t *fills);

/* ---- the steps either side of align(): postalign.c ---------------------- */
typedef struct { int32_t start, stop; } orc_index_pair_t; /* index_pair_t, GPU/f5c.h:184-188 */
typedef struct {                                          /* event_alignment_t, GPU/f5c.h:191-207 */
  char ref_kmer[7];
  int32_t ref_position;
  int32_t read_idx;
  int32_t event_idx;
  uint8_t rc; /* bool */
  char model_kmer[7];
  char hmm_state;
} orc_event_alignment_t;
typedef struct { float scale, shift, var, log_var; } orc_scalings_t; /* GPU/f5c.h:151-167 */

/* f5c postalign(), GPU/f5cmisc.h:51-52: AlignedPairs -> base-to-event map,
 * events per base and the event alignment records.  Returns the number of
 * records. */
int32_t orc_postalign(orc_event_alignment_t *alignment, orc_index_pair_t *base_to_event_map,
                      double *events_per_base, const char *sequence, int32_t n_kmers,
                      const orc_pair_t *pairs, int32_t n_pairs);

/* f5c recalibrate_model(), GPU/f5cmisc.h:53-54, with the event means in
 * place of the event table.  Returns 1 when the scalings were re-estimated. */
int orc_recalibrate_model(const orc_model_t *model, const float *ev_mean, orc_scalings_t *sc,
                          const orc_event_alignment_t *alignment, int32_t n_alignment,
                          int scale_var);

/* f5c estimate_scalings_using_mom(), GPU/f5cmisc.h:47. */
orc_scalings_t orc_estimate_scalings_mom(const char *sequence, int32_t sequence_len,
                                         const orc_model_t *model, const float *ev_mean,
                                         int32_t n_events);

/* ---- the step before: getevents.c ---------------------------------------- */
typedef struct { /* event_t, GPU/f5c.h:113-121 (24 bytes with tail padding) */
  uint64_t start;
  float length, mean, stdv;
} orc_event_t;

/* Prefix sums of a signal and of its (float) squares, n + 1 entries each. */
void orc_signal_prefix(const float *x, int64_t n, double *S, double *Q);
/* The windowed t-statistic of width w over the prefix sums, n entries. */
void orc_tstat(const double *S, const double *Q, int64_t n, int32_t w, float *t);
/* The short/long peak detector: boundaries in emission order, returns their
 * count (peaks has room for n). */
int64_t orc_detect_peaks(const float *t_short, const float *t_long, int64_t n, int64_t *peaks);
/* f5c getevents(), GPU/f5cmisc.h:44: the events of n raw samples.  Returns
 * the event count, minus the count when it exceeds cap, 0 when allocation
 * fails. */
int64_t orc_getevents(const float *raw, int64_t n, orc_event_t *out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
