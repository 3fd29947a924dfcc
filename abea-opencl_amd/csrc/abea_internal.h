// Shared between the HIP kernels and the host orchestration of libabea.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "abea.h"

namespace abea {

constexpr int kW = ABEA_BANDWIDTH;  // band width, GPU/f5c.h:30
constexpr int kSlots = 128;          // k-indexed circular band slots per wave (2 per lane)
constexpr int kRowBytes = 128;       // trace bytes per band row (one byte per slot)

// Per-read metadata, one 64-byte record per read (read-order index).
struct alignas(16) ReadMeta {
  int64_t trace_off;  // byte offset of the read's trace rows (n_bands * 128 B)
  int64_t km_off;     // offset (in k-mer parameter records) of the read's k-mers
  int64_t ev_off;     // offset of the read's events (= batch ev_ptr)
  int64_t pair_off;   // offset of the read's pair slot
  int64_t mv_off;     // offset (uint32 words) of the read's band-move bits
  int32_t n_e;        // events
  int32_t n_k;        // k-mers = read_len - 5 (<= 0 for reads that are skipped)
  double lp_stay;     // per-read transition log-probabilities (host glibc log/exp)
  double lp_step;
};
static_assert(sizeof(ReadMeta) == 64, "ReadMeta must stay 64 bytes");

// Staged per-k-mer emission parameters written by the pre kernel.
struct alignas(16) KmerParam {
  float mean;   // scale * level_mean + shift (f32 multiply then add)
  float stdv;   // level_stdv
  float c;      // -0.918938f - level_log_stdv
  int32_t rank; // 6-mer rank (for the model-gradient scatter)
};

struct KernelArgs {
  const ReadMeta *meta;
  const int32_t *queue;      // read indices in processing order
  int32_t n_queue;
  int32_t *counter;          // persistent work-queue head
  const char *seq;
  const int64_t *seq_ptr;
  const abea_event_t *events;
  const float *ev_mean_in;   // SoA event means supplied by the caller (or null)
  const abea_scalings_t *scalings;
  const abea_model_t *model;
  KmerParam *km;             // staged k-mer parameters
  float *ev;                 // staged event means (SoA)
  float *endcol;             // band cell (e, n_k - 1) per event
  uint8_t *trace;            // uint8 trace rows
  uint32_t *moves;           // 1 bit per band (1 = right)
  const uint32_t *replay;    // band moves to replay, or null
  uint8_t *rev_from;         // traceback from-codes in traceback order (scratch)
  double lp_skip;
  double lp_trim;
  // outputs
  abea_output_t out;
};

// Kernel launchers (abea_kernels.hip).
hipError_t launch_pre(const KernelArgs &a, int q_begin, int q_count, hipStream_t s);
hipError_t launch_core(const KernelArgs &a, int q_begin, int q_count, bool persistent,
                       int waves_per_block, int grid_blocks, hipStream_t s);
hipError_t launch_post(const KernelArgs &a, int q_begin, int q_count, int rows, hipStream_t s);
hipError_t launch_adjoint(const KernelArgs &a, const float *grad_score, float *grad_scalings,
                          double *grad_model, hipStream_t s);
hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t s);
int core_max_waves_per_block();

}  // namespace abea
