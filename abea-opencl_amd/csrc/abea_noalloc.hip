// abea_align_noalloc(): f5c's per-read align() for callers that cannot
// allocate device memory (serving loops and harnesses that forbid hipMalloc
// after start-up).  No counterpart in the reference.
//
// Every buffer the per-read forward needs (the read's bases, event records,
// scalings and model, the staged k-mer records and event means, the 2-bit
// trace, the band moves and K checkpoints, the traceback segment codes and
// the pair outputs) is carved out of one file-scope device array of
// kScratchBytes, which each GPU the library runs on holds until the process
// ends.  Calls on one device are serialised by a mutex, because they share
// that device's array; the array keeps the last call's contents, and nothing
// in it needs clearing: the walk writes every trace, move and checkpoint word
// that the traceback reads, and the traceback writes its segment codes
// before reading them.  The kernels are the plain form's (one core and one
// post launch), so the results are those of abea_align().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "abea_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_dev_mu[kMaxDevices];       // one call at a time per device
void *g_scratch_ptr[kMaxDevices] = {};  // the array's address on each device (set once)

int32_t noalloc_fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int32_t noalloc_fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  abea::set_last_error(buf);
  return -1;
}

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of one read's buffers inside the scratch array (the plan's
// per-read layout for a batch of one read).
struct Layout {
  size_t meta, queue, seq_ptr, scal, model, ends, counts, seq, events, km, km_rank, ev, trace,
      moves, kck, seg, pairs, end;
};

Layout layout(int64_t L, int64_t n_e, int64_t n_k) {
  const int64_t bands = n_e + n_k + 2;
  // staged arrays: the read's entries with kStagePad of padding on both
  // sides, then the slack the band walk's unclamped prefetch may read
  const int64_t tail = n_e + n_k + abea::kPad;
  const int64_t kmc = n_k + 2 * abea::kStagePad + tail;
  const int64_t evc = n_e + 2 * abea::kStagePad + tail;
  const int64_t tr = (bands + abea::kTileBands - 1) / abea::kTileBands * abea::kTileWords;
  const int64_t mv = (bands + 31) / 32;
  Layout y{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = up256(o + bytes);
    return at;
  };
  y.meta = take(sizeof(abea::ReadMeta));
  y.queue = take(sizeof(int32_t));
  y.seq_ptr = take(sizeof(int64_t));
  y.scal = take(sizeof(abea_scalings_t));
  y.model = take(sizeof(abea_model_t) * ABEA_NUM_KMER);
  y.ends = take(sizeof(abea::EndCell));
  y.counts = take(2 * sizeof(int32_t));  // n_pairs, path_len
  y.seq = take((size_t)L);
  y.events = take(sizeof(abea_event_t) * (size_t)n_e);
  y.km = take(sizeof(abea::KmerParam) * (size_t)kmc);
  y.km_rank = take(sizeof(uint16_t) * (size_t)kmc);
  y.ev = take(sizeof(float) * (size_t)evc);
  y.trace = take(sizeof(uint32_t) * (size_t)tr);
  y.moves = take(sizeof(uint32_t) * (size_t)mv);
  y.kck = take(sizeof(int32_t) * (size_t)mv);
  y.seg = take((size_t)((bands + 1024 + 3) & ~(int64_t)3));
  y.pairs = take(sizeof(abea_aligned_pair_t) * (size_t)(n_e + n_k));
  y.end = o;
  return y;
}

// The device address of this device's scratch array, looked up once per
// device with that device current.
int32_t scratch_base(int dev, unsigned char **base) {
  if (!g_scratch_ptr[dev]) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
      return noalloc_fail("abea_align_noalloc: hipGetDeviceProperties(%d) failed", dev);
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return noalloc_fail("abea_align_noalloc: device %d is %s; libabea is built for gfx950",
                          dev, prop.gcnArchName);
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return noalloc_fail("abea_align_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                          hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *base = (unsigned char *)g_scratch_ptr[dev];
  return 0;
}

#define NA_TRY(call)                                                                      \
  do {                                                                                    \
    const hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                               \
      (void)hipGetLastError();                                                            \
      return noalloc_fail("abea_align_noalloc: %s failed: %s", #call, hipGetErrorString(e_)); \
    }                                                                                     \
  } while (0)

}  // namespace

extern "C" int32_t abea_align_noalloc(abea_aligned_pair_t *out_2, char *sequence,
                                      int32_t sequence_len, abea_event_table events,
                                      abea_model_t *models, abea_scalings_t scaling,
                                      float sample_rate) {
  (void)sample_rate;  // unused by the reference too, GPU/align_cuda.cl:84-88
  if (!out_2 || !sequence || !models || (events.n && !events.event) || sequence_len < 0)
    return noalloc_fail("abea_align_noalloc: null argument");
  if (events.n > (size_t)0x7fffffff)
    return noalloc_fail("abea_align_noalloc: %zu events (need < 2^31)", events.n);
  const int32_t n_e = (int32_t)events.n;
  const int32_t n_k = sequence_len - ABEA_KMER_SIZE + 1;
  if (n_k <= 0 || n_e <= 0) return 0;  // nothing to align: a failed read, as align() gives
  const Layout y = layout(sequence_len, n_e, n_k);
  if (y.end > kScratchBytes)
    return noalloc_fail(
        "abea_align_noalloc: a read of %d bases and %d events needs %zu bytes of device "
        "scratch; the static array holds %zu (use abea_align for larger reads)",
        sequence_len, n_e, y.end, kScratchBytes);
  int dev = 0;
  NA_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return noalloc_fail("abea_align_noalloc: device %d outside 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lk(g_dev_mu[dev]);
  unsigned char *base = nullptr;
  if (scratch_base(dev, &base)) return -1;

  abea::ReadMeta m{};
  m.trace_off = 0;
  m.km_off = abea::kStagePad;
  m.sev_off = abea::kStagePad;
  m.ev_off = 0;
  m.pair_off = 0;
  m.mv_off = 0;
  m.seg_off = 0;
  m.bmap_off = 0;
  m.n_e = n_e;
  m.n_k = n_k;
  abea::transition_constants(n_e, n_k, &m.lp_stay, &m.lp_step);
  const int32_t queue0 = 0;
  const int64_t seq_ptr0 = 0;

  // the inputs, on the null stream (no stream of our own is created)
  const hipStream_t s = nullptr;
  NA_TRY(hipMemcpyAsync(base + y.meta, &m, sizeof m, hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(base + y.queue, &queue0, sizeof queue0, hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(base + y.seq_ptr, &seq_ptr0, sizeof seq_ptr0, hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(base + y.scal, &scaling, sizeof scaling, hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(base + y.model, models, sizeof(abea_model_t) * ABEA_NUM_KMER,
                        hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(base + y.seq, sequence, (size_t)sequence_len, hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(base + y.events, events.event, sizeof(abea_event_t) * (size_t)n_e,
                        hipMemcpyHostToDevice, s));

  abea::KernelArgs a{};
  a.meta = (const abea::ReadMeta *)(base + y.meta);
  a.queue = (const int32_t *)(base + y.queue);
  a.n_queue = 1;
  a.n_prio = 0;
  a.counter = nullptr;  // not persistent: the queue head is never read
  a.seq = (const char *)(base + y.seq);
  a.seq_ptr = (const int64_t *)(base + y.seq_ptr);
  a.events = (const abea_event_t *)(base + y.events);
  a.ev_mean_in = nullptr;
  a.scalings = (const abea_scalings_t *)(base + y.scal);
  a.model = (const abea_model_t *)(base + y.model);
  a.km = (abea::KmerParam *)(base + y.km);
  a.km_rank = (uint16_t *)(base + y.km_rank);
  a.ev = (float *)(base + y.ev);
  a.ends = (abea::EndCell *)(base + y.ends);
  a.trace = (uint32_t *)(base + y.trace);
  a.moves = (uint32_t *)(base + y.moves);
  a.kck = (int32_t *)(base + y.kck);
  a.replay = nullptr;
  a.seg = base + y.seg;
  a.lp_skip = log(1e-10);
  a.lp_trim = log(0.01);
  int32_t *counts = (int32_t *)(base + y.counts);
  a.out.pairs = (abea_aligned_pair_t *)(base + y.pairs);
  a.out.path_from = nullptr;
  a.out.n_pairs = counts;
  a.out.path_len = counts + 1;
  a.out.score = nullptr;
  a.out.avg_log_emission = nullptr;
  constexpr int kOverlap = 128;  // traceback overlap of the plain form's default tuning
  NA_TRY(abea::launch_core(a, 0, 1, false, 1, 1, 0, kOverlap, s));
  NA_TRY(abea::launch_post(a, 0, 1, kOverlap, s));

  int32_t c[2] = {0, 0};
  NA_TRY(hipMemcpyAsync(c, counts, sizeof c, hipMemcpyDeviceToHost, s));
  NA_TRY(hipStreamSynchronize(s));
  int32_t np = c[0];
  if (np < 0 || (int64_t)np > (int64_t)n_e + n_k)
    return noalloc_fail("abea_align_noalloc: the kernels returned %d pairs", np);
  if ((int64_t)np > 2 * (int64_t)n_e) return 0;  // would overrun f5c's 2*n_events buffer
  if (np) {
    NA_TRY(hipMemcpyAsync(out_2, base + y.pairs, sizeof(abea_aligned_pair_t) * (size_t)np,
                          hipMemcpyDeviceToHost, s));
    NA_TRY(hipStreamSynchronize(s));
  }
  return np;
}
