// Shared between the HIP kernels and the host orchestration of libabea.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "abea.h"

namespace abea {

constexpr int kW = ABEA_BANDWIDTH;  // band width, GPU/f5c.h:30
constexpr int kSlots = 128;          // k-indexed circular band slots per wave (2 per lane)
constexpr int kTileBands = 8;        // bands per trace tile
constexpr int kTileWords = 64;       // one 32-bit word per lane per tile: 8 bands x 2 slots x 2 bits
constexpr int kPad = 256;            // elements of slack after staged arrays (prefetch reads past a read)

// Per-read metadata, one 96-byte record per read (read-order index).
struct alignas(16) ReadMeta {
  int64_t trace_off;  // word offset of the read's 2-bit trace tiles (ceil(n_bands/8) * 64 words)
  int64_t km_off;     // offset (in k-mer parameter records) of the read's k-mers
  int64_t ev_off;     // offset of the read's events (= batch ev_ptr)
  int64_t pair_off;   // offset of the read's pair slot
  int64_t mv_off;     // offset (uint32 words) of the read's band-move bits and K checkpoints
  int64_t seg_off;    // 4-byte aligned byte offset of the read's traceback segment scratch
  int32_t n_e;        // events
  int32_t n_k;        // k-mers = read_len - 5 (<= 0 for reads that are skipped)
  double lp_stay;     // per-read transition log-probabilities (host glibc log/exp)
  double lp_step;
  int64_t bmap_off;   // offset of the read's base-to-event map entries (prefix of max(n_k, 0))
  int64_t sev_off;    // offset of the read's staged event means (kStagePad entries of padding
                      // before and after them; km_off likewise for the k-mer records)
  int64_t reserved;
};
static_assert(sizeof(ReadMeta) == 96, "ReadMeta must stay 96 bytes");

// Padding around each read's staged event means and k-mer records.  Padding
// events are kPadEvent and padding k-mers {0, 1, 1, -inf}: every emission
// that uses one is -inf, so band cells outside the matrix compute to -inf in
// the walk without masks.  A pad of 128 covers the walk's gathers (events down
// to E - 126 and up to E + 16 with E <= n_e + 111; k-mers from K >= -120 to
// K + 127 with K <= n_k).
constexpr int kStagePad = 128;
constexpr float kPadEvent = 1e30f;  // (x - m)/s squared overflows float for any stdv s < 5e10

// Staged per-k-mer emission parameters, written by the core wave of the read.
struct alignas(16) KmerParam {
  float mean;   // scale * level_mean + shift (f32 multiply then add)
  float stdv;   // level_stdv
  float rstdv;  // 1 / level_stdv, correctly rounded
  float c;      // -0.918938f - level_log_stdv
};

// Result of the end search over the last k-mer's column (per read).
struct EndCell {
  float score;  // best (float) end score, -inf when no end cell is in the band
  int32_t e;    // its event (the first maximum)
};

struct KernelArgs {
  const ReadMeta *meta;
  const int32_t *queue;      // read indices in processing order
  int32_t n_queue;
  int32_t n_prio;            // queue entries [0, n_prio) run at raised wave priority
  int32_t *counter;          // persistent work-queue head
  const char *seq;
  const int64_t *seq_ptr;
  const abea_event_t *events;
  const float *ev_mean_in;   // SoA event means supplied by the caller (or null)
  const abea_scalings_t *scalings;
  const abea_model_t *model;
  KmerParam *km;             // staged k-mer parameters
  uint16_t *km_rank;         // 6-mer rank per staged k-mer (model-gradient scatter)
  float *ev;                 // staged event means (SoA)
  EndCell *ends;             // per read: best end score and its event (the end search)
  uint32_t *trace;           // 2-bit from-codes in 8-band tiles
  uint32_t *moves;           // 1 bit per band (1 = right)
  int32_t *kck;              // K after every 32nd band (band 32w+31), per move word
  const uint32_t *replay;    // band moves to replay, or null
  uint8_t *seg;              // traceback segment scratch (from-codes, one byte each)
  double lp_skip;
  double lp_trim;
  // outputs
  abea_output_t out;
};

// Inputs and outputs of the post-align kernels (abea_postalign.hip): f5c's
// postalign + recalibrate_model from AlignedPairs, one wave per read.
struct PostArgs {
  int32_t n_reads;
  const ReadMeta *meta;          // n_e, n_k, ev_off, pair_off (pairs and records), bmap_off
  const int32_t *order;          // read processed by wave w: order[w] (null: w)
  const char *seq;
  const int64_t *seq_ptr;
  const abea_event_t *events;    // event records, or
  const float *ev_mean;          // SoA event means (one of the two)
  const abea_model_t *model;
  const abea_scalings_t *scal_in;
  const abea_aligned_pair_t *pairs;
  const int32_t *n_pairs;        // per read: f5c n_event_align_pairs
  abea_event_alignment_t *rec;   // optional
  abea_index_pair_t *bmap;       // optional
  int32_t *n_rec;
  double *epb;                   // optional
  abea_scalings_t *scal_out;
  int32_t *flag;
  const int2 *items;             // walk work items: (read, segment of kSegPairs pairs)
  int32_t n_items;
  int32_t *pstat;                // per read: the walk found a bad path (the finish kernel
                                 // reads and clears it; zero before the first call)
  float4 *terms;                 // recalibration scratch: one float4 per k-mer (bmap_off layout)
  int32_t recalibrate;           // 0: postalign only (no event means needed)
  int32_t scale_var;
};

// recalibrate_model from event alignment records (one read, one wave).
struct RecalArgs {
  const abea_event_alignment_t *rec;
  int32_t n_rec;
  const abea_event_t *events;
  int32_t n_events;
  const abea_model_t *model;
  abea_scalings_t *scal;         // in/out
  int32_t *status;               // 1 recalibrated, 0 not, -1 bad record
  int32_t scale_var;
};

// estimate_scalings_using_mom, one wave per read.
struct MomArgs {
  int32_t n_reads;
  const int32_t *order;          // reads in the order the waves take them (null: input order)
  const ReadMeta *meta;          // n_e, n_k, ev_off
  const char *seq;
  const int64_t *seq_ptr;
  const abea_event_t *events;
  const float *ev_mean;
  const abea_model_t *model;
  abea_scalings_t *out;
};

// Event detection (getevents), one read per raw signal.  Samples are cut
// into kGeChunk-sample chunks for the prefix sums and the events, and into
// detector chunks of GeArgs::det_chunk samples for the peak detector (with
// the t-statistics), chosen per call between kGeDetChunkMin and
// kGeDetChunkMax (kGeDetChunk when set by hand).
constexpr int kGeChunk = 2048;
constexpr int kGeDetChunk = 512;
constexpr int kGeDetChunkMin = 256, kGeDetChunkMax = 4096;
constexpr int kGeOverlap = 64;  // samples a detector chunk's speculative run starts early
// resident 256-lane detector blocks per CU (waves per SIMD): the detector
// kernel's occupancy bound, and the unit the chunk size is chosen for
constexpr int kGeDetBlocksPerCU = 4;
struct GeDet {
  int32_t pos;        // -1: no peak yet
  float val;          // running minimum, or the peak's value
  int32_t valid;
  int32_t masked_to;  // samples i <= masked_to are skipped
};
struct GeState {
  GeDet d[2];  // short, long
};
// per prefix chunk: its sums and what decides whether a read's sums are
// exact in any order
struct GeChunkSum {
  double s, q;        // sum of samples, of their float squares
  double abs_s, abs_q;  // sums of magnitudes (upper bounds of every partial sum)
  int32_t gx, gq;     // smallest ulp exponent of a sample / square (1000: all zero)
  int32_t bad;        // an inf or NaN
  int32_t hx;         // largest ulp exponent of a nonzero sample (-1000: none)
};
struct GeArgs {
  int32_t n_reads;
  int32_t overlap;
  int32_t det_chunk;          // samples per detector chunk
  int32_t pad_;
  int64_t n_chunks;           // prefix chunks
  int64_t n_det;              // detector chunks
  const int64_t *raw_off;     // n_reads + 1: read r is raw[raw_off[r] .. raw_off[r+1])
  const float *raw;
  double *S, *Q;              // prefix sums of serial reads, read r at raw_off[r] + r (n + 1)
  const int64_t *chunk_off;   // n_reads + 1: read r's prefix chunks
  const int32_t *chunk_read;  // the read of each prefix chunk
  GeChunkSum *csum;           // per prefix chunk
  double *carry;              // per prefix chunk: S and Q before it (2 doubles)
  int32_t *exact;             // per read: 1 when the parallel sums are exact, 2 when also
                              // every sample is below 2^60 in magnitude and every sample
                              // and float square is a multiple of 2^-125 (tame)
  const int64_t *det_off;     // n_reads + 1: read r's detector chunks
  const int32_t *det_read;    // the read of each detector chunk
  int32_t *peaks;             // det_chunk boundaries per detector chunk, in emission order
  int32_t *n_peaks;           // per detector chunk
  GeState *entry, *exit;      // per detector chunk: state before its first / after its last sample
  int32_t *ev_base;           // per detector chunk: the read's event ending at its first boundary
  int32_t *prev_last;         // per detector chunk: the read's last boundary before it (0: none)
  int32_t *first_bad;         // per read: first detector chunk whose entry is not its
                              // predecessor's exit (INT32_MAX: none)
  int32_t *reruns;            // per read: detector chunks re-run from the true state
  abea_event_t *events;       // read r at raw_off[r] + r (room for n + 1)
  int64_t *n_events;          // per read
};
hipError_t launch_getevents(const GeArgs &a, hipStream_t s);

// Per-read transition constants exactly as the CPU reference computes them
// (FPGA/code/device/align_swi.cl:233-250): double p_stay, glibc log/exp (host).
void transition_constants(int32_t n_e, int32_t n_k, double *lp_stay, double *lp_step);
// Sets the calling thread's abea_last_error() message.
void set_last_error(const char *msg);

// flag bit for a read whose pairs are not an alignment path (kernel guard)
constexpr int32_t kBadPath = 0x100;

hipError_t launch_postalign(const PostArgs &a, hipStream_t s);
int postalign_segment_pairs();
hipError_t launch_recalibrate_records(const RecalArgs &a, hipStream_t s);
hipError_t launch_mom(const MomArgs &a, hipStream_t s);

// Kernel launchers (abea_kernels.hip).
// Clears the persistent core's queue head (before each batched core launch).
hipError_t launch_queue_reset(const KernelArgs &a, hipStream_t s);
// fuse != 0: the core wave also does its read's traceback; otherwise the
// post kernel does.  overlap: rows each traceback segment starts above itself.
hipError_t launch_core(const KernelArgs &a, int q_begin, int q_count, bool persistent,
                       int waves_per_block, int grid_blocks, int fuse, int overlap, hipStream_t s);
hipError_t launch_post(const KernelArgs &a, int q_begin, int q_count, int overlap, hipStream_t s);
// Copies read r's n_pairs[r] pairs from its slots to dst + coff[r] (the
// host-buffer drop-ins' device-to-host compaction).
hipError_t launch_compact_pairs(const ReadMeta *meta, const abea_aligned_pair_t *pairs,
                                const int32_t *n_pairs, const int64_t *coff,
                                abea_aligned_pair_t *dst, int n_reads, hipStream_t s);
// grad_model: [4096][3] doubles (level_mean, level_stdv, level_log_stdv).
// partial holds n_blocks slices of 2 * 4096 * 3 fixed-point limbs (one per
// workgroup, n_blocks = adjoint_blocks()); special (4096 * 3 doubles) takes
// the non-finite or huge terms.
int adjoint_blocks(int n_reads, int n_cu);
hipError_t launch_adjoint(const KernelArgs &a, const float *grad_score, float *grad_scalings,
                          double *grad_model, unsigned long long *partial, int n_blocks,
                          double *special, hipStream_t s);
// Per-read state of the refinement loop's start selection.
struct RefineState {
  abea_scalings_t given;  // the scalings the loop started from
  float2 given_step;      // their Newton step's (scale, shift)
  float given_score;      // their best-path score
  int32_t pad;
};
static_assert(sizeof(RefineState) == 32, "RefineState must stay 32 bytes");
// step: device counter selecting the start-selection phase (null: a plain
// Newton step); start: the alternative start (method-of-moments scalings).
hipError_t launch_refine(const KernelArgs &a, abea_scalings_t *scal_out, float damping,
                         const int32_t *step, const abea_scalings_t *start, RefineState *state,
                         hipStream_t s);
// set_to >= 0 stores it, otherwise increments the counter.
hipError_t launch_refine_step_counter(int32_t *step, int32_t set_to, hipStream_t s);
int core_max_waves_per_block();

}  // namespace abea
