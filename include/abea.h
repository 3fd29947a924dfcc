/*
 * abea.h -- C ABI of the MI355X (gfx950) Adaptive Banded Event Alignment
 * engine.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Drop-in boundary (reference interfaces replaced):
 *   abea_align()        f5c per-read align(),  GPU/f5cmisc.h:48-50
 *   abea_align_db()     f5c batched align_cuda(core_t*, db_t*),
 *                       GPU/f5cmisc.h:99-102; the reference's OpenCL
 *                       analogue is align_ocl(), GPU/align_cuda.c:391-943
 *   abea_plan_*()       the same batch, device-resident, for callers that
 *                       keep their data in HBM (the torch binding, benches)
 *
 * Semantics follow f5c's CPU align(): AlignedPair lists identical to it,
 * band W = 100, k = 6, 4096-row pore model.
 */
#ifndef ABEA_H
#define ABEA_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ABEA_BANDWIDTH 100 /* ALN_BANDWIDTH, GPU/f5c.h:30 */
#define ABEA_KMER_SIZE 6   /* KMER_SIZE, GPU/f5c.h:26 */
#define ABEA_NUM_KMER 4096 /* NUM_KMER, GPU/f5c.h:27 */

/* ---- f5c data types, byte-identical to GPU/f5c.h:112-181 (x86-64) ---- */
typedef struct {
  uint64_t start;
  float length;
  float mean;
  float stdv;
} abea_event_t; /* event_t, GPU/f5c.h:113-121 (24 bytes with tail padding) */

typedef struct {
  size_t n;
  size_t start;
  size_t end;
  abea_event_t *event;
} abea_event_table; /* event_table, GPU/f5c.h:124-130 */

typedef struct {
  float level_mean;
  float level_stdv;
  float level_log_stdv;
} abea_model_t; /* model_t with CACHED_LOG, GPU/f5c.h:133-148 */

typedef struct {
  float scale;
  float shift;
  float var;
  float log_var;
} abea_scalings_t; /* scalings_t, GPU/f5c.h:151-167 */

typedef struct {
  int32_t ref_pos;
  int32_t read_pos;
} abea_aligned_pair_t; /* AlignedPair, GPU/f5c.h:177-181 */

typedef struct {
  int32_t start;
  int32_t stop; /* inclusive */
} abea_index_pair_t; /* index_pair_t, GPU/f5c.h:184-188 */

typedef struct {
  char ref_kmer[ABEA_KMER_SIZE + 1];
  int32_t ref_position;
  int32_t read_idx;
  int32_t event_idx;
  bool rc;
  char model_kmer[ABEA_KMER_SIZE + 1];
  char hmm_state; /* 'M' or 'E' */
} abea_event_alignment_t; /* event_alignment_t, GPU/f5c.h:191-207 (32 bytes) */

/* read status bits, db_t.read_stat_flag (GPU/f5c.h:49-52) */
#define ABEA_FAILED_CALIBRATION 0x001
#define ABEA_FAILED_ALIGNMENT 0x002
#define ABEA_FAILED_QUALITY_CHK 0x004
/* f5c's bound on the recalibrated var (read fails calibration above it) */
#define ABEA_MIN_CALIBRATION_VAR 2.5f
/* f5c's stay filter: a calibrated read with more events per base than this
 * fails the quality check */
#define ABEA_MAX_EVENTS_PER_BASE 5.0

typedef enum {
  ABEA_OK = 0,
  ABEA_ERR_INVALID = 1,  /* bad argument */
  ABEA_ERR_HIP = 2,      /* a HIP runtime call failed */
  ABEA_ERR_NOMEM = 3,    /* device or host allocation failed */
  ABEA_ERR_CAPACITY = 4, /* output capacity too small */
  ABEA_ERR_NO_DEVICE = 5 /* no usable gfx950 device */
} abea_status_t;

/* Human-readable message for the last error raised on this thread. */
const char *abea_last_error(void);

/* ---------------------------------------------------------------------- */
/* (1) f5c align(): one read, host buffers.  GPU/f5cmisc.h:48-50.
 * Returns the number of pairs written to out_2 (0 = QC failed, as f5c).
 * out_2 must hold 2 * events.n pairs (f5c's allocation,
 * GPU/align_cuda.c:496); a path that would not fit is reported as a failed
 * read (0) instead of overrunning.  sample_rate is unused, as in the
 * reference (GPU/align_cuda.cl:84-88).  Threads may call it concurrently:
 * calls that overlap in time on the current device are gathered into one
 * batched forward on a pooled context (abea_host_contexts()), with the same
 * results as one call at a time; a call may wait up to 200 us for others to
 * join its batch.  Returns -1 on a device error (abea_last_error() tells
 * which). */
int32_t abea_align(abea_aligned_pair_t *out_2, char *sequence, int32_t sequence_len,
                   abea_event_table events, abea_model_t *models, abea_scalings_t scaling,
                   float sample_rate);

/* (1') abea_align() for callers that cannot allocate device memory: a
 * serving loop or harness that forbids hipMalloc after start-up.  No
 * counterpart in the reference.  Same arguments, results and return values
 * as abea_align(), but it allocates nothing: every device buffer the read
 * needs comes from one static 64 MiB device array that each GPU the library
 * runs on holds until the process ends, and it runs on the null stream of
 * the current device.  A read needs about 90 bytes per band (n_events +
 * n_kmers + 2 bands), so reads up to about 700 k bands (a 250 kb read at
 * 1.6 events per base) fit; a larger read returns -1 with
 * abea_last_error() saying so.  Calls on one device share its array and
 * run one at a time (a mutex). */
int32_t abea_align_noalloc(abea_aligned_pair_t *out_2, char *sequence, int32_t sequence_len,
                           abea_event_table events, abea_model_t *models,
                           abea_scalings_t scaling, float sample_rate);

/* (2) f5c align_cuda(core, db): a batch in host buffers.
 * GPU/f5cmisc.h:99-102 / align_ocl GPU/align_cuda.c:391-943.
 * Reads read[i] (read_len[i] bases), et[i], scalings[i] and model[4096];
 * writes n_event_align_pairs[i] and event_align_pairs[i], which is
 * malloc()ed here per read (the caller frees it), as align_ocl does
 * (GPU/align_cuda.c:926-931).  Uses the batched persistent form. */
abea_status_t abea_align_db(int32_t n_bam_rec, char **read, const int32_t *read_len,
                            const abea_event_table *et, const abea_model_t *model,
                            const abea_scalings_t *scalings,
                            abea_aligned_pair_t **event_align_pairs,
                            int32_t *n_event_align_pairs);

/* Phases of the calling thread's last abea_align() or abea_align_db() call,
 * in ms: [0] flatten (host), [1] host-to-device copies of the inputs, [2]
 * planning (per-read metadata, sort, its upload), [3] kernels, [4]
 * device-to-host copies ([1]-[4] hipEvent-timed on the call's stream), [5]
 * unflatten (host), [6] the whole call.  The reference's host times the
 * same phases (GPU/align_cuda.c:59-74, 378-385). */
void abea_host_last_times(double times_ms[7]);

/* The host-buffer entry points draw device contexts (a stream and device
 * buffers) from a process-wide pool: a call takes one for the current device
 * and returns it when it ends, so calls from any number of threads (f5c's
 * align() workers, GPU/f5c.h:57-59) hold at most as many contexts as ran at
 * once.  Returns how many contexts the pool has created. */
int32_t abea_host_contexts(void);

/* ---------------------------------------------------------------------- */
/* (1b) The f5c steps either side of align(), per read, host buffers.  Each
 * runs on a pooled context of the current device (one small launch); use the
 * batched or plan forms below for many reads.  On a device error they set
 * abea_last_error() and return: NaN scalings (MOM), -1 (postalign), false
 * (recalibrate_model). */

/* f5c estimate_scalings_using_mom(), GPU/f5cmisc.h:47: method-of-moments
 * shift and scale from the event means and the read's k-mer levels; var and
 * log_var are 0, as in the scalings f5c passes to align(). */
abea_scalings_t abea_estimate_scalings_using_mom(char *sequence, int32_t sequence_len,
                                                 abea_model_t *pore_model, abea_event_table et);

/* f5c postalign(), GPU/f5cmisc.h:51-52: an aligned read's AlignedPairs
 * (event_alignment, n_events of them, as align() returned) become the
 * base-to-event map (n_kmers entries), events_per_base and one
 * event_alignment_t per event on the path (at most n_events records).
 * The pairs must form an alignment path (ascending, each step moving to the
 * next event, the next k-mer or both, k-mers in [0, n_kmers)); other input
 * is rejected with -1.  Returns the number of records. */
int32_t abea_postalign(abea_event_alignment_t *alignment, abea_index_pair_t *base_to_event_map,
                       double *events_per_base, char *sequence, int32_t n_kmers,
                       abea_aligned_pair_t *event_alignment, int32_t n_events);

/* f5c recalibrate_model(), GPU/f5cmisc.h:53-54: weighted least-squares
 * shift/scale (and, with scale_var, var and log_var) from the 'M' records;
 * needs at least 200 of them, otherwise returns false and leaves *scallings
 * unchanged. */
bool abea_recalibrate_model(abea_model_t *pore_model, abea_event_table et,
                            abea_scalings_t *scallings,
                            const abea_event_alignment_t *alignment_output,
                            int32_t num_alignments, bool scale_var);

/* Batched forms over f5c's db_t arrays (GPU/f5c.h:291-348), one launch per
 * batch.  abea_estimate_scalings_db writes scalings[i] for every read.
 * abea_postalign_db runs f5c's per-read flow after align_cuda(): for reads
 * with n_event_align_pairs[i] > 0 it malloc()s event_alignment[i]
 * (n_event_align_pairs[i] records) and base_to_event_map[i] (read_len[i] - 5
 * entries), runs postalign and recalibrate_model(scale_var = true) and sets
 * read_stat_flag[i] to ABEA_FAILED_CALIBRATION when recalibration did not
 * run or var > ABEA_MIN_CALIBRATION_VAR, else to ABEA_FAILED_QUALITY_CHK
 * when events_per_base[i] > ABEA_MAX_EVENTS_PER_BASE (f5c's filter of reads
 * with too many stays), else 0; reads align() failed get
 * ABEA_FAILED_ALIGNMENT, NULL arrays, n_event_alignment[i] = 0 and
 * events_per_base[i] = 0.  scalings[] is updated in place, as f5c does.
 * The caller frees the malloc()ed arrays. */
abea_status_t abea_estimate_scalings_db(int32_t n_bam_rec, char **read, const int32_t *read_len,
                                        const abea_event_table *et, const abea_model_t *model,
                                        abea_scalings_t *scalings);
abea_status_t abea_postalign_db(int32_t n_bam_rec, char **read, const int32_t *read_len,
                                const abea_event_table *et, const abea_model_t *model,
                                abea_aligned_pair_t *const *event_align_pairs,
                                const int32_t *n_event_align_pairs, abea_scalings_t *scalings,
                                abea_event_alignment_t **event_alignment,
                                int32_t *n_event_alignment, double *events_per_base,
                                abea_index_pair_t **base_to_event_map, int32_t *read_stat_flag);

/* ---------------------------------------------------------------------- */
/* (1c) f5c getevents(), GPU/f5cmisc.h:44: event detection on one read's raw
 * signal in picoamperes (f5c's rawptr), by scrappie's DNA detector as f5c
 * ports it (the event types are scrappie's, GPU/f5c.h:112-130): two windowed
 * t-statistics (w = 3 and 6) over double prefix sums, the short/long peak
 * detector (thresholds 1.4 and 9.0, peak height 0.2), and events counted
 * as scrappie's create_events counts them: m boundaries give m events, from
 * 0 to the first boundary, between consecutive boundaries, and the last from
 * the second-last boundary to the end (the last boundary ends no event).
 * Returns the event table with a malloc()ed event array that the caller
 * frees, start = 0 and end = n, as getevents does; a signal with fewer than
 * two boundaries is one event (the C code reads before its boundary array
 * there; nsample = 0: no events).  On error: n = 0, event = NULL,
 * abea_last_error() set.  nsample < 2^31. */
abea_event_table abea_getevents(size_t nsample, float *rawptr);

/* getevents over a batch in host buffers (signal i: rawptr[i], nsample[i]);
 * et[i] as abea_getevents returns it. */
abea_status_t abea_getevents_db(int32_t n_reads, const size_t *nsample, float *const *rawptr,
                                abea_event_table *et);

/* ---------------------------------------------------------------------- */
/* (3) Device-resident batches.  All pointers in abea_batch_t and
 * abea_output_t are device pointers except the two *_host arrays. */

typedef struct abea_ctx abea_ctx_t;
typedef struct abea_plan abea_plan_t;

/* One context per device per thread of use.  stream 0 = a new stream owned
 * by the context; otherwise the caller's hipStream_t (cast to void*). */
abea_status_t abea_ctx_create(abea_ctx_t **ctx, int device, void *stream);
void abea_ctx_destroy(abea_ctx_t *ctx);
void *abea_ctx_stream(abea_ctx_t *ctx);

/* getevents on signals already in device memory (new; GPU/f5cmisc.h:44 per
 * read).  Read i is raw[raw_off_host[i] .. raw_off_host[i+1]) (raw device,
 * raw_off_host a host array of n_reads + 1 non-decreasing offsets, each
 * read < 2^31 samples); its events go to events[raw_off_host[i] + i ...]
 * (device, room for nsample + 1 each) and their count to n_events[i]
 * (device).  Enqueued on the context's stream; the context keeps the
 * scratch (about 28 bytes per sample). */
abea_status_t abea_getevents_device(abea_ctx_t *ctx, int32_t n_reads,
                                    const int64_t *raw_off_host, const float *raw,
                                    abea_event_t *events, int64_t *n_events);
/* Samples each detector chunk's speculative run starts before the chunk
 * (default 64, 0..512).  Results do not depend on it, only how many
 * chunks are re-run from the true state; abea_ctx_getevents_reruns() gives
 * that count for the last abea_getevents_device() call (synchronizes). */
abea_status_t abea_ctx_set_getevents_overlap(abea_ctx_t *ctx, int32_t samples);
int64_t abea_ctx_getevents_reruns(abea_ctx_t *ctx);
/* Samples per detector chunk (one lane walks each chunk): 0 (the default)
 * picks it per call, 256..4096, so that the detector's lanes fill whole
 * rounds of the GPU's resident waves; otherwise 256..4096 fixed.  Results do
 * not depend on it.  abea_ctx_getevents_chunk() gives the last call's. */
abea_status_t abea_ctx_set_getevents_chunk(abea_ctx_t *ctx, int32_t samples);
int32_t abea_ctx_getevents_chunk(abea_ctx_t *ctx);

typedef struct {
  int32_t n_reads;
  const char *seq;               /* bases, CSR by seq_ptr (no NUL needed) */
  const int64_t *seq_ptr;        /* [n_reads] */
  const int32_t *read_len;       /* [n_reads] */
  const abea_event_t *events;    /* [sum n_events] event_t records, CSR by ev_ptr */
  const float *ev_mean;          /* alternative to events: means only (SoA); one of the two */
  const int64_t *ev_ptr;         /* [n_reads] */
  const int32_t *n_events;       /* [n_reads] */
  const abea_scalings_t *scalings; /* [n_reads] */
  const abea_model_t *model;     /* [4096] */
  const int32_t *read_len_host;  /* host copies, used for planning */
  const int32_t *n_events_host;
  const int64_t *seq_ptr_host;
  const int64_t *ev_ptr_host;
} abea_batch_t;

/* Per-read output slots: read i owns pairs [off[i], off[i] + n_events[i] +
 * max(read_len[i]-5, 0)), where off is the exclusive prefix sum of those
 * capacities in read order (abea_pair_offsets()); abea_pair_capacity()
 * gives the total. */
typedef struct {
  abea_aligned_pair_t *pairs;      /* ascending pairs of the traced path (kept for QC-failed reads) */
  uint8_t *path_from;              /* optional: 0 = D, 1 = U, 2 = L per pair */
  int32_t *n_pairs;                /* f5c's n_event_align_pairs: path length, 0 when QC fails */
  int32_t *path_len;               /* optional: path length before QC */
  float *score;                    /* optional: best-path score (f5c float semantics) */
  double *avg_log_emission;        /* optional: QC statistic */
} abea_output_t;

int64_t abea_pair_capacity(int32_t n_reads, const int32_t *read_len_host,
                           const int32_t *n_events_host);
void abea_pair_offsets(int32_t n_reads, const int32_t *read_len_host,
                       const int32_t *n_events_host, int64_t *offsets);

enum {
  ABEA_PLAN_PLAIN = 0,   /* one pre/core/post launch set per read, in read order */
  ABEA_PLAN_BATCHED = 1, /* length-sorted work queue, persistent core, graph replay */
};

/* Plan a batch: uploads per-read metadata and sizes the workspace (trace
 * matrix, staged k-mer parameters, ...).  The batch's device buffers must
 * stay valid (and may be updated in place, e.g. new scalings) for the plan's
 * lifetime. */
abea_status_t abea_plan_create(abea_ctx_t *ctx, const abea_batch_t *batch, int form,
                               abea_plan_t **plan);
void abea_plan_destroy(abea_plan_t *plan);

/* Run pre -> core -> post on the context's stream (asynchronous). */
abea_status_t abea_plan_forward(abea_plan_t *plan, const abea_output_t *out);

/* Record the band moves of the last forward in moves[n_words]: per read, in
 * read order, ceil(n_bands / 32) words (reads with no k-mers or no events
 * take none); bit b % 32 of word b / 32 is band b's move (1 = right). */
int64_t abea_plan_move_words(const abea_plan_t *plan);
abea_status_t abea_plan_get_moves(abea_plan_t *plan, uint32_t *moves_host);
/* Hold the band fixed: later forwards replay these moves instead of
 * Suzuki's rule (NULL restores adaptive banding). */
abea_status_t abea_plan_set_replay(abea_plan_t *plan, const uint32_t *moves_host);

/* Adjoint of the summed weighted best-path score.  Needs the pairs and
 * path_from of a forward on the same plan.  grad_score[n_reads] (device,
 * f32) weights each read.  Outputs (device; overwritten, no zeroing needed):
 *   grad_scalings[n_reads*3]  d/d(shift, scale, var), f32 (var column is 0:
 *                             the reference emission ignores var,
 *                             GPU/align_cuda.cl:95)
 *   grad_model[4096*3]        per k-mer rank, the partial derivatives in
 *                             (level_mean, level_stdv, level_log_stdv), f64:
 *                             the forward reads the cached level_log_stdv as
 *                             an input of its own (GPU/align_cuda.cl:517-518),
 *                             so d/d level_stdv holds it fixed; with
 *                             level_log_stdv = log(level_stdv) the total
 *                             derivative is column 1 + column 2 / level_stdv.
 * Deterministic: the k-mer sums are exact fixed-point sums of the per-run
 * terms (integer adds, whose order does not matter), so repeated calls give
 * identical bits. */
abea_status_t abea_plan_adjoint(abea_plan_t *plan, const abea_output_t *fwd,
                                const float *grad_score, float *grad_scalings,
                                double *grad_model);

/* Tuning knobs chosen by the autotuner (or set by hand). */
typedef struct {
  int32_t core_waves_per_block; /* 1, 2 or 4 waves per workgroup */
  int32_t core_waves_per_cu;    /* resident waves per CU for the persistent grid */
  int32_t traceback_overlap;    /* rows above its segment where each traceback lane starts
                                   following the trace (32..1024); affects time, not results */
  int32_t use_graph;            /* replay pre->core->post as one hipGraph */
  int32_t fuse_traceback;       /* 0: separate post kernel; 1: each core wave traces its
                                   own read back */
} abea_tuning_t;

abea_status_t abea_plan_get_tuning(const abea_plan_t *plan, abea_tuning_t *t);
abea_status_t abea_plan_set_tuning(abea_plan_t *plan, const abea_tuning_t *t);
/* Time the candidate configurations on this plan (runs forwards: each
 * candidate reps times, the three fastest reps + 3 times more, a candidate's
 * time being its fastest run) and keep the fastest; best_ms receives its
 * time. */
abea_status_t abea_plan_autotune(abea_plan_t *plan, const abea_output_t *out, int reps,
                                 double *best_ms);

/* Scaling refinement (new; no reference counterpart: f5c re-estimates
 * scalings by least squares after alignment, recalibrate_model(),
 * GPU/f5cmisc.h:53-54).  One Gauss-Newton step on each read's (scale,
 * shift) along the path of the last forward on this plan: on a fixed path
 * the score is quadratic in (scale, shift), so the step moves to that
 * path's optimum (times damping).  scalings_out (device, n_reads records)
 * receives the new scalings (var, log_var copied); it may be the batch's
 * own scalings array, updated in place.  Needs path_from and path_len in
 * fwd.  Asynchronous on the context stream. */
abea_status_t abea_plan_refine_step(abea_plan_t *plan, const abea_output_t *fwd,
                                    abea_scalings_t *scalings_out, float damping);
/* The refinement loop: steps x (forward + refine step) with the batch's
 * scalings updated in place, each step replayed as one captured graph;
 * returns when done.  flags = ABEA_REFINE_MOM_START first picks each read's
 * start: the loop runs one more forward, at the method-of-moments scalings
 * (abea_plan_estimate_scalings), and each read continues from whichever of
 * its given and its MOM scalings gave the higher best-path score (needs
 * score in out).  A read whose alignment is lost under its given scalings
 * has a path that says nothing about them, and Newton steps along it do not
 * recover the read.  ms_per_step (optional) receives the loop's device time
 * divided by steps (the start's forward and the MOM estimate included). */
#define ABEA_REFINE_MOM_START 1
abea_status_t abea_plan_refine(abea_plan_t *plan, const abea_output_t *out, int steps,
                               float damping, int flags, double *ms_per_step);

/* Post-align outputs (device pointers).  Records of read i start at its pair
 * offset (abea_pair_offsets()); its base-to-event map entries at
 * abea_kmer_offsets(), the exclusive prefix sum of max(read_len - 5, 0). */
typedef struct {
  abea_event_alignment_t *alignment;     /* optional: event alignment records */
  abea_index_pair_t *base_to_event_map;  /* optional */
  int32_t *n_alignment;                  /* f5c n_event_alignment (0 when align's QC failed) */
  double *events_per_base;               /* optional */
  abea_scalings_t *scalings;             /* recalibrated scalings; a read that is not
                                            recalibrated keeps its input scalings.  May be
                                            the batch's own scalings array. */
  int32_t *read_stat_flag;               /* ABEA_FAILED_* bits */
} abea_postalign_output_t;

void abea_kmer_offsets(int32_t n_reads, const int32_t *read_len_host, int64_t *offsets);

/* f5c's flow after align() (postalign, then recalibrate_model with
 * scale_var, then the read status) for every read of the batch, from the
 * pairs and n_pairs of a forward (fwd).  Asynchronous on the context stream.
 * No reference batched counterpart: f5c runs these per read on host threads
 * after align_cuda() (GPU/f5cmisc.h:51-54). */
abea_status_t abea_plan_postalign(abea_plan_t *plan, const abea_output_t *fwd,
                                  const abea_postalign_output_t *out, int scale_var);
/* f5c estimate_scalings_using_mom() for every read of the batch into out[]
 * (device, n_reads records; may be the batch's own scalings array).
 * Asynchronous on the context stream. */
abea_status_t abea_plan_estimate_scalings(abea_plan_t *plan, abea_scalings_t *out);

/* Per-kernel device time (ms) of the last forward/adjoint, hipEvent-timed
 * on the plan's stream: [0]=queue reset (batched form; the core stages each
 * read itself) [1]=core [2]=post [3]=adjoint.  Where the kernels cannot be
 * told apart, [1] holds the whole forward and [0], [2] are -1: the plain
 * form (one launch pair per read) and graph replays of the batched form
 * (timed by events around the launch; the queue reset is a one-wave kernel
 * of a few microseconds). */
abea_status_t abea_plan_last_times(const abea_plan_t *plan, double times_ms[4]);
/* Bytes of device workspace held by the plan. */
int64_t abea_plan_workspace_bytes(const abea_plan_t *plan);

#ifdef __cplusplus
}
#endif
#endif /* ABEA_H */
