// Host side of libabea: contexts, batch plans, graph replay, the autotuner
// and the f5c drop-in entry points declared in include/abea.h.
//
// Orchestration replaces the reference's align_ocl() (GPU/align_cuda.c:
// 391-943): flatten -> buffers -> pre/core/post -> readback.  Differences by
// design: device buffers are planned once per batch and reused, reads are
// queued longest-first for a persistent core that also stages each read (the
// reference's pre kernel), and the launches are replayed as one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "abea_internal.h"

using abea::KernelArgs;
using abea::KmerParam;
using abea::ReadMeta;

namespace {

thread_local std::string g_err;

abea_status_t fail(abea_status_t st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

// A failed call is reported once: HIP's last-error state is cleared, so a
// later launch's hipGetLastError() does not report it again.
#define HIP_TRY(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (void)hipGetLastError();                                                         \
      return fail(ABEA_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                 \
    }                                                                                  \
  } while (0)

// Pinned host staging (hipHostMalloc), grown as needed and kept: the
// host-buffer drop-ins flatten their inputs into it and copy results back
// into it, so that both copies run at the link's rate.
struct PinBuf {
  void *p = nullptr;
  size_t n = 0;
  ~PinBuf() {
    if (p) (void)hipHostFree(p);
  }
  abea_status_t ensure(size_t bytes) {
    if (bytes <= n && p) return ABEA_OK;
    if (p) {
      (void)hipHostFree(p);
      p = nullptr;
      n = 0;
    }
    // a quarter more than asked, so that slowly growing batches reallocate rarely
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(ABEA_ERR_NOMEM, "hipHostMalloc(%zu bytes) failed: %s", want,
                  hipGetErrorString(e));
    }
    n = want;
    return ABEA_OK;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  abea_status_t ensure(size_t bytes) {
    if (bytes <= n && p) return ABEA_OK;
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      n = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(ABEA_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
    }
    n = want;
    return ABEA_OK;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace

namespace abea {

void transition_constants(int32_t n_e, int32_t n_k, double *lp_stay, double *lp_step) {
  double events_per_kmer = (double)n_e / n_k;
  double p_stay = 1 - (1 / (events_per_kmer + 1));
  double lp_skip = log(1e-10);
  *lp_stay = log(p_stay);
  *lp_step = log(1.0 - exp(lp_skip) - exp(*lp_stay));
}

void set_last_error(const char *msg) { g_err = msg; }

}  // namespace abea

using abea::transition_constants;

// Event-detection scratch of a context (grows, never shrinks).
struct GeScratch {
  // the call's chunk tables (read offsets, prefix and detector chunk offsets,
  // then the read of each chunk), staged in pinned memory and copied to the
  // device as one block
  PinBuf tab_h;
  DevBuf tab;
  hipEvent_t tab_copied = nullptr;  // the last call's copy out of tab_h
  DevBuf S, Q, ev_base, prev_last, csum, carry, exact, peaks, n_peaks, entry, exit, first_bad,
      reruns;
  int32_t n_reads = 0;
  ~GeScratch() {
    if (tab_copied) (void)hipEventDestroy(tab_copied);
  }
};

struct abea_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int n_cu = 256;
  std::string arch;
  int32_t ge_overlap = abea::kGeOverlap;
  int32_t ge_chunk = 0;  // detector chunk samples; 0: chosen per call
  int32_t ge_last_chunk = abea::kGeDetChunk;
  GeScratch ge;
};

struct abea_plan {
  abea_ctx *ctx = nullptr;
  int form = ABEA_PLAN_BATCHED;
  int32_t n = 0;
  abea_batch_t batch{};
  std::vector<ReadMeta> meta_h;
  std::vector<int32_t> order_h;
  int64_t total_pairs = 0, trace_words = 0, move_words = 0, km_count = 0, ev_count = 0;
  int64_t total_kmers = 0;  // base-to-event map entries (sum of max(n_k, 0))
  int64_t n_bands_total = 0, seg_bytes = 0, tail_pad = 0;
  int32_t n_prio = 0;  // leading queue entries run at raised wave priority
  DevBuf meta, queue, counter, km, km_rank, ev, ends, trace, moves, kck, replay, seg;
  // adjoint: fixed-point k-mer gradient bins (allocated by the first adjoint)
  DevBuf gacc, gspecial;
  // refinement loop: start scalings, per-read start state, step counter
  DevBuf rstart, rstate, rstep;
  // postalign scratch, allocated by the first postalign call: per-k-mer
  // recalibration terms, per-read walk status (zeroed once, then kept zero by
  // the finish kernel) and the walk's work items
  DevBuf terms, pstat, items;
  int32_t n_items = 0;
  bool post_ready = false;
  std::vector<int2> items_h;
  bool replay_on = false;
  double lp_skip = 0, lp_trim = 0;
  abea_tuning_t tune{1, 24, 128, 1, 1};
  // graph cache
  hipGraphExec_t gexec = nullptr;
  abea_output_t gout{};
  abea_tuning_t gtune{};
  bool greplay = false;
  hipEvent_t evt[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double times[4] = {0, 0, 0, 0};
  bool have_times = false;

  ~abea_plan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (auto &e : evt)
      if (e) (void)hipEventDestroy(e);
  }
};

extern "C" {

const char *abea_last_error(void) { return g_err.c_str(); }

int64_t abea_pair_capacity(int32_t n_reads, const int32_t *read_len_host,
                           const int32_t *n_events_host) {
  int64_t t = 0;
  for (int32_t i = 0; i < n_reads; ++i)
    t += (int64_t)n_events_host[i] + std::max<int64_t>((int64_t)read_len_host[i] - 5, 0);
  return t;
}

void abea_pair_offsets(int32_t n_reads, const int32_t *read_len_host, const int32_t *n_events_host,
                       int64_t *offsets) {
  int64_t t = 0;
  for (int32_t i = 0; i < n_reads; ++i) {
    offsets[i] = t;
    t += (int64_t)n_events_host[i] + std::max<int64_t>((int64_t)read_len_host[i] - 5, 0);
  }
}

abea_status_t abea_ctx_create(abea_ctx_t **out, int device, void *stream) {
  if (!out) return fail(ABEA_ERR_INVALID, "abea_ctx_create: null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
    return fail(ABEA_ERR_NO_DEVICE, "abea_ctx_create: no HIP device %d (found %d)", device, count);
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  std::string arch = prop.gcnArchName;
  if (arch.rfind("gfx950", 0) != 0)
    return fail(ABEA_ERR_NO_DEVICE, "abea_ctx_create: device %d is %s; libabea is built for gfx950",
                device, arch.c_str());
  abea_ctx *c = new abea_ctx();
  c->device = device;
  c->arch = arch;
  c->n_cu = prop.multiProcessorCount;
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(ABEA_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  *out = c;
  return ABEA_OK;
}

void abea_ctx_destroy(abea_ctx_t *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);  // work on the context's scratch
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

void *abea_ctx_stream(abea_ctx_t *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"

namespace {

// Detector chunk size for a call.  The detector kernel is bound by VALU issue
// and runs kGeDetBlocksPerCU 4-wave blocks per CU, one lane per chunk, each lane walking
// its chunk plus the lead-in; the kernel takes about (rounds of resident
// blocks) x (chunk + lead-in) sample steps.  The chunk is chosen to fill the
// last round, among multiples of 32 from kGeDetChunkMin to kGeDetChunkMax
// (ties to the larger chunk: fewer chunks to check and re-run).
int32_t pick_det_chunk(int32_t n, const int64_t *raw_off_h, int n_cu, int32_t overlap) {
  const int64_t slots = (int64_t)std::max(n_cu, 1) * abea::kGeDetBlocksPerCU;  // resident blocks
  int32_t best = abea::kGeDetChunk;
  double best_cost = 1e300;
  for (int32_t D = abea::kGeDetChunkMin; D <= abea::kGeDetChunkMax; D += 32) {
    int64_t chunks = 0, longest = 0;
    for (int32_t r = 0; r < n; ++r) {
      const int64_t len = raw_off_h[r + 1] - raw_off_h[r];
      chunks += (len + D - 1) / D;
      longest = std::max(longest, len);
    }
    const int64_t blocks = (chunks + 255) / 256;
    const int64_t rounds = std::max<int64_t>((blocks + slots - 1) / slots, 1);
    const double cost = (double)rounds * (double)(std::min<int64_t>(D, longest) + overlap);
    if (cost <= best_cost) {
      best_cost = cost;
      best = D;
    }
  }
  return best;
}

// getevents over n signals in device memory (see abea_getevents_device).
abea_status_t ge_enqueue(abea_ctx *ctx, int32_t n, const int64_t *raw_off_h, const float *raw,
                         abea_event_t *events, int64_t *n_events) {
  if (n < 0 || (n > 0 && (!raw_off_h || !events || !n_events)))
    return fail(ABEA_ERR_INVALID, "getevents: null argument or negative read count");
  if (n == 0) return ABEA_OK;
  if (raw_off_h[0] < 0) return fail(ABEA_ERR_INVALID, "getevents: negative raw offset");
  GeScratch &g = ctx->ge;
  for (int32_t r = 0; r < n; ++r) {
    const int64_t len = raw_off_h[r + 1] - raw_off_h[r];
    if (len < 0 || len >= ((int64_t)1 << 31))
      return fail(ABEA_ERR_INVALID, "getevents: read %d has %lld samples (need 0 .. 2^31-1)", r,
                  (long long)len);
  }
  const int32_t D = ctx->ge_chunk > 0 ? ctx->ge_chunk
                                      : pick_det_chunk(n, raw_off_h, ctx->n_cu, ctx->ge_overlap);
  ctx->ge_last_chunk = D;
  if (!raw && raw_off_h[n] > 0) return fail(ABEA_ERR_INVALID, "getevents: null raw signal");
  int64_t nc = 0, nd = 0;
  for (int32_t r = 0; r < n; ++r) {
    const int64_t len = raw_off_h[r + 1] - raw_off_h[r];
    nc += (len + abea::kGeChunk - 1) / abea::kGeChunk;
    nd += (len + D - 1) / D;
  }
  const int64_t extent = raw_off_h[n];
  const size_t nc1 = (size_t)std::max<int64_t>(nc, 1), nd1 = (size_t)std::max<int64_t>(nd, 1);
  // the tables: three int64 arrays of n + 1, then the int32 chunk-to-read maps
  const size_t off_bytes = sizeof(int64_t) * (size_t)(n + 1);
  const size_t tab_bytes = 3 * off_bytes + sizeof(int32_t) * (size_t)(nc + nd);
  hipStream_t s = ctx->stream;
  abea_status_t st;
  if (!g.tab_copied) HIP_TRY(hipEventCreateWithFlags(&g.tab_copied, hipEventDisableTiming));
  // the previous call's copy may still read the staging block
  HIP_TRY(hipEventSynchronize(g.tab_copied));
  if ((st = g.tab_h.ensure(tab_bytes)) || (st = g.tab.ensure(tab_bytes)) ||
      (st = g.S.ensure(sizeof(double) * (size_t)(extent + n + 1))) ||
      (st = g.Q.ensure(sizeof(double) * (size_t)(extent + n + 1))) ||
      (st = g.ev_base.ensure(sizeof(int32_t) * nd1)) ||
      (st = g.prev_last.ensure(sizeof(int32_t) * nd1)) ||
      (st = g.csum.ensure(sizeof(abea::GeChunkSum) * nc1)) ||
      (st = g.carry.ensure(2 * sizeof(double) * nc1)) ||
      (st = g.exact.ensure(sizeof(int32_t) * (size_t)n)) ||
      (st = g.peaks.ensure(sizeof(int32_t) * nd1 * (size_t)D)) ||
      (st = g.n_peaks.ensure(sizeof(int32_t) * nd1)) ||
      (st = g.entry.ensure(sizeof(abea::GeState) * nd1)) ||
      (st = g.exit.ensure(sizeof(abea::GeState) * nd1)) ||
      (st = g.first_bad.ensure(sizeof(int32_t) * (size_t)n)) ||
      (st = g.reruns.ensure(sizeof(int32_t) * (size_t)n)))
    return st;
  int64_t *raw_off_t = g.tab_h.as<int64_t>();
  int64_t *chunk_off_t = raw_off_t + (n + 1);
  int64_t *det_off_t = chunk_off_t + (n + 1);
  int32_t *chunk_read_t = reinterpret_cast<int32_t *>(det_off_t + (n + 1));
  int32_t *det_read_t = chunk_read_t + nc;
  std::memcpy(raw_off_t, raw_off_h, off_bytes);
  chunk_off_t[0] = 0;
  det_off_t[0] = 0;
  for (int32_t r = 0; r < n; ++r) {
    const int64_t len = raw_off_h[r + 1] - raw_off_h[r];
    const int64_t nch = (len + abea::kGeChunk - 1) / abea::kGeChunk;
    const int64_t ndc = (len + D - 1) / D;
    for (int64_t k = 0; k < nch; ++k) chunk_read_t[chunk_off_t[r] + k] = r;
    for (int64_t k = 0; k < ndc; ++k) det_read_t[det_off_t[r] + k] = r;
    chunk_off_t[r + 1] = chunk_off_t[r] + nch;
    det_off_t[r + 1] = det_off_t[r] + ndc;
  }
  HIP_TRY(hipMemcpyAsync(g.tab.p, g.tab_h.p, tab_bytes, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(g.tab_copied, s));
  const int64_t *raw_off_d = g.tab.as<int64_t>();
  const int64_t *chunk_off_d = raw_off_d + (n + 1);
  const int64_t *det_off_d = chunk_off_d + (n + 1);
  const int32_t *chunk_read_d = reinterpret_cast<const int32_t *>(det_off_d + (n + 1));
  const int32_t *det_read_d = chunk_read_d + nc;
  abea::GeArgs a{};
  a.n_reads = n;
  a.overlap = ctx->ge_overlap;
  a.det_chunk = D;
  a.n_chunks = nc;
  a.n_det = nd;
  a.raw_off = raw_off_d;
  a.raw = raw;
  a.S = g.S.as<double>();
  a.Q = g.Q.as<double>();
  a.ev_base = g.ev_base.as<int32_t>();
  a.prev_last = g.prev_last.as<int32_t>();
  a.chunk_off = chunk_off_d;
  a.chunk_read = chunk_read_d;
  a.csum = g.csum.as<abea::GeChunkSum>();
  a.carry = g.carry.as<double>();
  a.exact = g.exact.as<int32_t>();
  a.det_off = det_off_d;
  a.det_read = det_read_d;
  a.peaks = g.peaks.as<int32_t>();
  a.n_peaks = g.n_peaks.as<int32_t>();
  a.entry = g.entry.as<abea::GeState>();
  a.exit = g.exit.as<abea::GeState>();
  a.first_bad = g.first_bad.as<int32_t>();
  a.reruns = g.reruns.as<int32_t>();
  a.events = events;
  a.n_events = n_events;
  HIP_TRY(abea::launch_getevents(a, s));
  g.n_reads = n;
  return ABEA_OK;
}

}  // namespace

extern "C" {

abea_status_t abea_getevents_device(abea_ctx_t *ctx, int32_t n_reads, const int64_t *raw_off_host,
                                    const float *raw, abea_event_t *events, int64_t *n_events) {
  if (!ctx) return fail(ABEA_ERR_INVALID, "abea_getevents_device: null context");
  HIP_TRY(hipSetDevice(ctx->device));
  return ge_enqueue(ctx, n_reads, raw_off_host, raw, events, n_events);
}

abea_status_t abea_ctx_set_getevents_overlap(abea_ctx_t *ctx, int32_t samples) {
  if (!ctx) return fail(ABEA_ERR_INVALID, "abea_ctx_set_getevents_overlap: null context");
  if (samples < 0 || samples > abea::kGeDetChunk)
    return fail(ABEA_ERR_INVALID, "getevents overlap %d outside 0..%d", samples,
                abea::kGeDetChunk);
  ctx->ge_overlap = samples;
  return ABEA_OK;
}

abea_status_t abea_ctx_set_getevents_chunk(abea_ctx_t *ctx, int32_t samples) {
  if (!ctx) return fail(ABEA_ERR_INVALID, "abea_ctx_set_getevents_chunk: null context");
  if (samples != 0 && (samples < abea::kGeDetChunkMin || samples > abea::kGeDetChunkMax))
    return fail(ABEA_ERR_INVALID, "getevents chunk %d: 0 or %d..%d", samples,
                abea::kGeDetChunkMin, abea::kGeDetChunkMax);
  ctx->ge_chunk = samples;
  return ABEA_OK;
}

int32_t abea_ctx_getevents_chunk(abea_ctx_t *ctx) { return ctx ? ctx->ge_last_chunk : 0; }

int64_t abea_ctx_getevents_reruns(abea_ctx_t *ctx) {
  if (!ctx || ctx->ge.n_reads <= 0) return 0;
  std::vector<int32_t> h((size_t)ctx->ge.n_reads);
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipMemcpyAsync(h.data(), ctx->ge.reruns.p, sizeof(int32_t) * h.size(),
                     hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    fail(ABEA_ERR_HIP, "abea_ctx_getevents_reruns: readback failed");
    return -1;
  }
  int64_t t = 0;
  for (int32_t v : h) t += v;
  return t;
}

}  // extern "C"

namespace {

// (Re)plan a batch on an existing plan object; buffers only grow.
abea_status_t plan_bind(abea_plan *p, const abea_batch_t *b, int form) {
  if (!b || b->n_reads < 0) return fail(ABEA_ERR_INVALID, "plan: bad batch");
  if (b->n_reads > 0 && (!b->read_len_host || !b->n_events_host || !b->ev_ptr_host ||
                         !b->seq || !b->seq_ptr || !b->scalings || !b->model ||
                         (!b->events && !b->ev_mean)))
    return fail(ABEA_ERR_INVALID, "plan: missing batch arrays");
  const int32_t n = b->n_reads;
  p->batch = *b;
  p->n = n;
  p->form = form;
  p->meta_h.assign(n, ReadMeta{});
  int64_t tr = 0, km = 0, pr = 0, mv = 0, evmax = 0, nb = 0, sg = 0, maxlen = 0, bm = 0, sev = 0;
  for (int32_t i = 0; i < n; ++i) {
    ReadMeta &m = p->meta_h[i];
    const int32_t n_e = b->n_events_host[i];
    const int32_t n_k = b->read_len_host[i] - ABEA_KMER_SIZE + 1;
    if (n_e < 0) return fail(ABEA_ERR_INVALID, "plan: read %d has %d events", i, n_e);
    m.n_e = n_e;
    m.n_k = n_k;
    m.ev_off = b->ev_ptr_host[i];
    evmax = std::max<int64_t>(evmax, m.ev_off + n_e);
    m.pair_off = pr;
    pr += (int64_t)n_e + std::max<int32_t>(n_k, 0);
    m.km_off = km + abea::kStagePad;
    m.sev_off = sev + abea::kStagePad;
    m.bmap_off = bm;
    bm += std::max<int32_t>(n_k, 0);
    m.trace_off = tr;
    m.mv_off = mv;
    m.seg_off = sg;
    if (n_k > 0 && n_e > 0) {
      const int64_t bands = (int64_t)n_e + n_k + 2;
      km += n_k + 2 * abea::kStagePad;
      sev += n_e + 2 * abea::kStagePad;
      tr += (bands + abea::kTileBands - 1) / abea::kTileBands * abea::kTileWords;
      mv += (bands + 31) / 32;
      // traceback segments: one code byte per row they cover; the last
      // segment ends at the path's end row, so the codes stay below
      // rows + 4 bytes (slack kept)
      sg += (bands + 1024 + 3) & ~(int64_t)3;
      nb += bands;
      maxlen = std::max<int64_t>(maxlen, (int64_t)n_e + n_k);
      transition_constants(n_e, n_k, &m.lp_stay, &m.lp_step);
    }
  }
  p->total_pairs = pr;
  p->total_kmers = bm;
  p->trace_words = tr;
  p->seg_bytes = sg;
  p->move_words = mv;
  p->km_count = km;
  p->ev_count = sev;
  p->n_bands_total = nb;
  // The core prefetches event E+1 and k-mer K+128 of the band walk without
  // clamping: E <= n_e + n_k + 50 and K + 128 <= n_e + n_k + 77 for any read,
  // so staged arrays carry this much slack past the last read.
  p->tail_pad = maxlen + abea::kPad;
  p->lp_skip = log(1e-10);
  p->lp_trim = log(0.01);

  // processing order: the plain form walks reads in input order; the batched
  // form queues them longest-first so the persistent core ends balanced
  p->order_h.resize(n);
  std::iota(p->order_h.begin(), p->order_h.end(), 0);
  if (form == ABEA_PLAN_BATCHED) {
    std::stable_sort(p->order_h.begin(), p->order_h.end(), [&](int32_t x, int32_t y) {
      const ReadMeta &a = p->meta_h[x], &c = p->meta_h[y];
      const int64_t ba = (int64_t)a.n_e + std::max(a.n_k, 0), bc = (int64_t)c.n_e + std::max(c.n_k, 0);
      return ba > bc;
    });
  }

  // reads within a factor of two of the longest set the critical path
  p->n_prio = 0;
  if (form == ABEA_PLAN_BATCHED && n > 0) {
    auto bands = [&](int32_t i) {
      const ReadMeta &m = p->meta_h[i];
      return (m.n_k > 0 && m.n_e > 0) ? (int64_t)m.n_e + m.n_k + 2 : 0;
    };
    const int64_t top = bands(p->order_h[0]);
    while (p->n_prio < n && 2 * bands(p->order_h[p->n_prio]) >= top && top > 0) ++p->n_prio;
  }

  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  const int64_t kmc = km + p->tail_pad, evc = sev + p->tail_pad;  // staged, padded per read
  const size_t need = (size_t)tr * 4 + (size_t)kmc * (sizeof(KmerParam) + 2) + (size_t)evc * 4 +
                      (size_t)mv * 12 + (size_t)sg + (size_t)n * (sizeof(ReadMeta) + 12);
  if (free_b && need > free_b + p->trace.n + p->km.n + p->km_rank.n + p->ev.n + p->ends.n +
                           p->moves.n + p->kck.n + p->replay.n + p->seg.n + p->meta.n)
    return fail(ABEA_ERR_NOMEM, "plan: workspace of %.2f GB exceeds free device memory (%.2f GB)",
                need / 1e9, free_b / 1e9);

  abea_status_t st;
  if ((st = p->meta.ensure(sizeof(ReadMeta) * (size_t)std::max(n, 1)))) return st;
  if ((st = p->queue.ensure(sizeof(int32_t) * (size_t)std::max(n, 1)))) return st;
  if ((st = p->counter.ensure(256))) return st;
  if ((st = p->km.ensure(sizeof(KmerParam) * (size_t)kmc))) return st;
  if ((st = p->km_rank.ensure(sizeof(uint16_t) * (size_t)kmc))) return st;
  if ((st = p->ev.ensure(sizeof(float) * (size_t)evc))) return st;
  if ((st = p->ends.ensure(sizeof(abea::EndCell) * (size_t)std::max(n, 1)))) return st;
  if ((st = p->trace.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(tr, 1)))) return st;
  if ((st = p->moves.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(mv, 1)))) return st;
  if ((st = p->kck.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(mv, 1)))) return st;
  if ((st = p->seg.ensure((size_t)std::max<int64_t>(sg, 4)))) return st;
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(p->meta.p, p->meta_h.data(), sizeof(ReadMeta) * n,
                           hipMemcpyHostToDevice, p->ctx->stream));
    HIP_TRY(hipMemcpyAsync(p->queue.p, p->order_h.data(), sizeof(int32_t) * n,
                           hipMemcpyHostToDevice, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  }
  p->replay_on = false;
  p->post_ready = false;
  if (p->gexec) {
    (void)hipGraphExecDestroy(p->gexec);
    p->gexec = nullptr;
  }
  p->tune.use_graph = form == ABEA_PLAN_BATCHED ? 1 : 0;
  // default residency: 16 waves per CU (four per SIMD), which keeps the
  // VALUs busy (97%) with the fewest cache lines in play; on the 16,384-read
  // workload 20, 24 and 32 waves per CU each ran about 3% slower
  p->tune.core_waves_per_cu = 16;
  return ABEA_OK;
}

KernelArgs make_args(const abea_plan *p, const abea_output_t *out) {
  KernelArgs a{};
  a.meta = p->meta.as<ReadMeta>();
  a.queue = p->queue.as<int32_t>();
  a.n_queue = p->n;
  a.n_prio = p->n_prio;
  a.counter = p->counter.as<int32_t>();
  a.seq = p->batch.seq;
  a.seq_ptr = p->batch.seq_ptr;
  a.events = p->batch.events;
  a.ev_mean_in = p->batch.ev_mean;
  a.scalings = p->batch.scalings;
  a.model = p->batch.model;
  a.km = p->km.as<KmerParam>();
  a.km_rank = p->km_rank.as<uint16_t>();
  a.ev = p->ev.as<float>();
  a.ends = p->ends.as<abea::EndCell>();
  a.trace = p->trace.as<uint32_t>();
  a.moves = p->moves.as<uint32_t>();
  a.kck = p->kck.as<int32_t>();
  a.replay = p->replay_on ? p->replay.as<uint32_t>() : nullptr;
  a.seg = p->seg.as<uint8_t>();
  a.lp_skip = p->lp_skip;
  a.lp_trim = p->lp_trim;
  if (out) a.out = *out;
  return a;
}

abea_status_t ensure_events(abea_plan *p) {
  for (auto &e : p->evt)
    if (!e) HIP_TRY(hipEventCreate(&e));
  return ABEA_OK;
}

// The launch sequence of one forward; evt non-null records per-kernel
// boundaries on the stream.
abea_status_t enqueue_forward(abea_plan *p, const KernelArgs &a, bool with_events) {
  hipStream_t s = p->ctx->stream;
  const abea_tuning_t &t = p->tune;
  if (with_events) HIP_TRY(hipEventRecord(p->evt[0], s));
  if (p->form == ABEA_PLAN_PLAIN) {
    // reference-shaped: one launch set per read, reads in input order
    for (int32_t i = 0; i < p->n; ++i) {
      HIP_TRY(abea::launch_core(a, i, 1, false, 1, 1, 0, t.traceback_overlap, s));
      HIP_TRY(abea::launch_post(a, i, 1, t.traceback_overlap, s));
    }
    if (with_events) {
      HIP_TRY(hipEventRecord(p->evt[1], s));
      HIP_TRY(hipEventRecord(p->evt[2], s));
      HIP_TRY(hipEventRecord(p->evt[3], s));
    }
    return ABEA_OK;
  }
  HIP_TRY(abea::launch_queue_reset(a, s));
  if (with_events) HIP_TRY(hipEventRecord(p->evt[1], s));
  const int wpb = t.core_waves_per_block;
  int grid = std::max(1, (p->ctx->n_cu * t.core_waves_per_cu) / wpb);
  grid = std::min(grid, std::max(1, (p->n + wpb - 1) / wpb));
  HIP_TRY(abea::launch_core(a, 0, p->n, true, wpb, grid, t.fuse_traceback, t.traceback_overlap, s));
  if (with_events) HIP_TRY(hipEventRecord(p->evt[2], s));
  if (!t.fuse_traceback) HIP_TRY(abea::launch_post(a, 0, p->n, t.traceback_overlap, s));
  if (with_events) HIP_TRY(hipEventRecord(p->evt[3], s));
  return ABEA_OK;
}

bool same_out(const abea_output_t &x, const abea_output_t &y) {
  return std::memcmp(&x, &y, sizeof(x)) == 0;
}
bool same_tune(const abea_tuning_t &x, const abea_tuning_t &y) {
  return std::memcmp(&x, &y, sizeof(x)) == 0;
}

abea_status_t check_out(const abea_plan *p, const abea_output_t *out) {
  if (!out || (p->n > 0 && (!out->pairs || !out->n_pairs)))
    return fail(ABEA_ERR_INVALID, "forward: output pairs and n_pairs are required");
  return ABEA_OK;
}

}  // namespace

extern "C" {

abea_status_t abea_plan_create(abea_ctx_t *ctx, const abea_batch_t *batch, int form,
                               abea_plan_t **plan) {
  if (!ctx || !plan) return fail(ABEA_ERR_INVALID, "abea_plan_create: null argument");
  if (form != ABEA_PLAN_PLAIN && form != ABEA_PLAN_BATCHED)
    return fail(ABEA_ERR_INVALID, "abea_plan_create: unknown form %d", form);
  HIP_TRY(hipSetDevice(ctx->device));
  abea_plan *p = new abea_plan();
  p->ctx = ctx;
  abea_status_t st = plan_bind(p, batch, form);
  if (st) {
    delete p;
    return st;
  }
  *plan = p;
  return ABEA_OK;
}

void abea_plan_destroy(abea_plan_t *p) { delete p; }

abea_status_t abea_plan_forward(abea_plan_t *p, const abea_output_t *out) {
  if (!p) return fail(ABEA_ERR_INVALID, "forward: null plan");
  abea_status_t st = check_out(p, out);
  if (st) return st;
  if (p->n == 0) return ABEA_OK;
  HIP_TRY(hipSetDevice(p->ctx->device));
  if ((st = ensure_events(p))) return st;
  const KernelArgs a = make_args(p, out);
  hipStream_t s = p->ctx->stream;
  p->have_times = false;
  if (p->tune.use_graph && p->form == ABEA_PLAN_BATCHED) {
    if (!p->gexec || !same_out(p->gout, *out) || !same_tune(p->gtune, p->tune) ||
        p->greplay != p->replay_on) {
      if (p->gexec) {
        (void)hipGraphExecDestroy(p->gexec);
        p->gexec = nullptr;
      }
      // The graph holds the kernels only.  Events recorded inside a graph
      // are timed unreliably (elapsed times a percent or so above the
      // stream's own, and refused outright when other events are recorded
      // on the stream between replays), so a replay is timed as a whole by
      // events around its launch.
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      abea_status_t cst = enqueue_forward(p, a, false);
      hipError_t ce = hipStreamEndCapture(s, &g);
      if (cst || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return cst ? cst : fail(ABEA_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
      }
      hipError_t ie = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        p->gexec = nullptr;
        return fail(ABEA_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
      }
      p->gout = *out;
      p->gtune = p->tune;
      p->greplay = p->replay_on;
    }
    HIP_TRY(hipEventRecord(p->evt[0], s));
    HIP_TRY(hipGraphLaunch(p->gexec, s));
    HIP_TRY(hipEventRecord(p->evt[3], s));
    p->have_times = true;
    return ABEA_OK;
  }
  st = enqueue_forward(p, a, true);
  if (!st) p->have_times = true;
  return st;
}

abea_status_t abea_plan_last_times(const abea_plan_t *pc, double times_ms[4]) {
  abea_plan *p = const_cast<abea_plan *>(pc);
  if (!p || !times_ms) return fail(ABEA_ERR_INVALID, "last_times: null argument");
  for (int i = 0; i < 4; ++i) times_ms[i] = -1.0;
  if (p->have_times && p->evt[0]) {
    HIP_TRY(hipEventSynchronize(p->evt[3]));
    float ms = 0;
    const bool split = !(p->tune.use_graph && p->form == ABEA_PLAN_BATCHED);
    if (split && p->form == ABEA_PLAN_BATCHED) {
      HIP_TRY(hipEventElapsedTime(&ms, p->evt[0], p->evt[1]));
      times_ms[0] = ms;
      HIP_TRY(hipEventElapsedTime(&ms, p->evt[1], p->evt[2]));
      times_ms[1] = ms;
      HIP_TRY(hipEventElapsedTime(&ms, p->evt[2], p->evt[3]));
      times_ms[2] = ms;
    } else {
      // whole forward only: the plain form interleaves the kernels, and a
      // graph replay is timed as a whole
      HIP_TRY(hipEventElapsedTime(&ms, p->evt[0], p->evt[3]));
      times_ms[1] = ms;
    }
  }
  times_ms[3] = p->times[3];
  return ABEA_OK;
}

int64_t abea_plan_workspace_bytes(const abea_plan_t *p) {
  if (!p) return 0;
  return (int64_t)(p->meta.n + p->queue.n + p->counter.n + p->km.n + p->km_rank.n + p->ev.n +
                   p->ends.n + p->trace.n + p->moves.n + p->kck.n + p->replay.n + p->seg.n +
                   p->terms.n + p->pstat.n + p->items.n + p->gacc.n + p->gspecial.n +
                   p->rstart.n + p->rstate.n + p->rstep.n);
}

int64_t abea_plan_move_words(const abea_plan_t *p) { return p ? p->move_words : 0; }

abea_status_t abea_plan_get_moves(abea_plan_t *p, uint32_t *moves_host) {
  if (!p || !moves_host) return fail(ABEA_ERR_INVALID, "get_moves: null argument");
  if (p->move_words == 0) return ABEA_OK;
  HIP_TRY(hipMemcpyAsync(moves_host, p->moves.p, sizeof(uint32_t) * p->move_words,
                         hipMemcpyDeviceToHost, p->ctx->stream));
  HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  return ABEA_OK;
}

abea_status_t abea_plan_set_replay(abea_plan_t *p, const uint32_t *moves_host) {
  if (!p) return fail(ABEA_ERR_INVALID, "set_replay: null plan");
  if (!moves_host) {
    p->replay_on = false;
    return ABEA_OK;
  }
  abea_status_t st = p->replay.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(p->move_words, 1));
  if (st) return st;
  if (p->move_words)
    HIP_TRY(hipMemcpyAsync(p->replay.p, moves_host, sizeof(uint32_t) * p->move_words,
                           hipMemcpyHostToDevice, p->ctx->stream));
  HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  p->replay_on = true;
  return ABEA_OK;
}

abea_status_t abea_plan_adjoint(abea_plan_t *p, const abea_output_t *fwd, const float *grad_score,
                                float *grad_scalings, double *grad_model) {
  if (!p || !fwd || !grad_score || !grad_scalings || !grad_model)
    return fail(ABEA_ERR_INVALID, "adjoint: null argument");
  if (!fwd->pairs || !fwd->path_from || !fwd->path_len)
    return fail(ABEA_ERR_INVALID, "adjoint: the forward must record pairs, path_from and path_len");
  HIP_TRY(hipSetDevice(p->ctx->device));
  abea_status_t st = ensure_events(p);
  if (st) return st;
  hipStream_t s = p->ctx->stream;
  const KernelArgs a = make_args(p, fwd);
  const int nb = abea::adjoint_blocks(p->n, p->ctx->n_cu);
  if ((st = p->gacc.ensure(sizeof(unsigned long long) * 6 * ABEA_NUM_KMER * (size_t)nb)) ||
      (st = p->gspecial.ensure(sizeof(double) * 3 * ABEA_NUM_KMER)))
    return st;
  HIP_TRY(hipEventRecord(p->evt[4], s));
  HIP_TRY(abea::launch_adjoint(a, grad_score, grad_scalings, grad_model,
                               p->gacc.as<unsigned long long>(), nb, p->gspecial.as<double>(),
                               s));
  HIP_TRY(hipEventRecord(p->evt[3], s));
  HIP_TRY(hipEventSynchronize(p->evt[3]));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, p->evt[4], p->evt[3]));
  p->times[3] = ms;
  p->have_times = false;
  return ABEA_OK;
}

abea_status_t abea_plan_refine_step(abea_plan_t *p, const abea_output_t *fwd,
                                    abea_scalings_t *scalings_out, float damping) {
  if (!p || !fwd || !scalings_out) return fail(ABEA_ERR_INVALID, "refine_step: null argument");
  if (!fwd->pairs || !fwd->path_from || !fwd->path_len)
    return fail(ABEA_ERR_INVALID, "refine_step: the forward must record pairs, path_from and path_len");
  if (p->n == 0) return ABEA_OK;
  HIP_TRY(hipSetDevice(p->ctx->device));
  const KernelArgs a = make_args(p, fwd);
  HIP_TRY(abea::launch_refine(a, scalings_out, damping, nullptr, nullptr, nullptr,
                              p->ctx->stream));
  return ABEA_OK;
}

void abea_kmer_offsets(int32_t n_reads, const int32_t *read_len_host, int64_t *offsets) {
  int64_t t = 0;
  for (int32_t i = 0; i < n_reads; ++i) {
    offsets[i] = t;
    t += std::max<int64_t>((int64_t)read_len_host[i] - ABEA_KMER_SIZE + 1, 0);
  }
}

abea_status_t abea_plan_postalign(abea_plan_t *p, const abea_output_t *fwd,
                                  const abea_postalign_output_t *out, int scale_var) {
  if (!p || !fwd || !out) return fail(ABEA_ERR_INVALID, "postalign: null argument");
  if (!fwd->pairs || !fwd->n_pairs)
    return fail(ABEA_ERR_INVALID, "postalign: the forward must record pairs and n_pairs");
  if (!out->n_alignment || !out->scalings || !out->read_stat_flag)
    return fail(ABEA_ERR_INVALID, "postalign: n_alignment, scalings and read_stat_flag are required");
  if (p->n == 0) return ABEA_OK;
  HIP_TRY(hipSetDevice(p->ctx->device));
  abea::PostArgs a{};
  a.n_reads = p->n;
  a.meta = p->meta.as<ReadMeta>();
  a.order = p->form == ABEA_PLAN_BATCHED ? p->queue.as<int32_t>() : nullptr;
  a.seq = p->batch.seq;
  a.seq_ptr = p->batch.seq_ptr;
  a.events = p->batch.events;  // records win over means, as in the staging
  a.ev_mean = p->batch.events ? nullptr : p->batch.ev_mean;
  a.model = p->batch.model;
  a.scal_in = p->batch.scalings;
  a.pairs = fwd->pairs;
  a.n_pairs = fwd->n_pairs;
  a.rec = out->alignment;
  a.bmap = out->base_to_event_map;
  a.n_rec = out->n_alignment;
  a.epb = out->events_per_base;
  a.scal_out = out->scalings;
  a.flag = out->read_stat_flag;
  if (!p->post_ready) {
    // walk work items in queue order, by pair capacity (the pair counts are
    // on the device); items past a read's last pair exit at once
    p->items_h.clear();
    const int seg = abea::postalign_segment_pairs();
    for (int32_t i = 0; i < p->n; ++i) {
      const int32_t r = p->order_h[i];
      const ReadMeta &m = p->meta_h[r];
      const int64_t cap = (int64_t)m.n_e + std::max(m.n_k, 0);
      for (int64_t q = 0; q * seg < cap; ++q) p->items_h.push_back(make_int2(r, (int)q));
    }
    p->n_items = (int32_t)p->items_h.size();
    abea_status_t st;
    if ((st = p->items.ensure(sizeof(int2) * std::max<size_t>(p->items_h.size(), 1)))) return st;
    if ((st = p->terms.ensure(sizeof(float4) * (size_t)std::max<int64_t>(p->total_kmers, 1))))
      return st;
    if ((st = p->pstat.ensure(sizeof(int32_t) * (size_t)std::max(p->n, 1)))) return st;
    if (p->n_items)
      HIP_TRY(hipMemcpyAsync(p->items.p, p->items_h.data(), sizeof(int2) * p->items_h.size(),
                             hipMemcpyHostToDevice, p->ctx->stream));
    HIP_TRY(hipMemsetAsync(p->pstat.p, 0, sizeof(int32_t) * (size_t)std::max(p->n, 1),
                           p->ctx->stream));
    p->post_ready = true;
  }
  a.items = p->items.as<int2>();
  a.n_items = p->n_items;
  a.pstat = p->pstat.as<int32_t>();
  a.terms = p->terms.as<float4>();
  a.recalibrate = 1;
  a.scale_var = scale_var ? 1 : 0;
  HIP_TRY(abea::launch_postalign(a, p->ctx->stream));
  return ABEA_OK;
}

abea_status_t abea_plan_estimate_scalings(abea_plan_t *p, abea_scalings_t *out) {
  if (!p || !out) return fail(ABEA_ERR_INVALID, "estimate_scalings: null argument");
  if (p->n == 0) return ABEA_OK;
  HIP_TRY(hipSetDevice(p->ctx->device));
  abea::MomArgs a{};
  a.n_reads = p->n;
  // the batched form's queue is longest first; the plain form keeps input order
  a.order = p->form == ABEA_PLAN_BATCHED ? p->queue.as<int32_t>() : nullptr;
  a.meta = p->meta.as<ReadMeta>();
  a.seq = p->batch.seq;
  a.seq_ptr = p->batch.seq_ptr;
  a.events = p->batch.events;  // records win over means, as in the staging
  a.ev_mean = p->batch.events ? nullptr : p->batch.ev_mean;
  a.model = p->batch.model;
  a.out = out;
  HIP_TRY(abea::launch_mom(a, p->ctx->stream));
  return ABEA_OK;
}

abea_status_t abea_plan_refine(abea_plan_t *p, const abea_output_t *out, int steps, float damping,
                               int flags, double *ms_per_step) {
  if (!p || !out) return fail(ABEA_ERR_INVALID, "refine: null argument");
  if (!out->pairs || !out->path_from || !out->path_len || !out->n_pairs)
    return fail(ABEA_ERR_INVALID, "refine: out must record pairs, path_from, path_len, n_pairs");
  if (steps < 0) return fail(ABEA_ERR_INVALID, "refine: negative step count");
  if (flags & ~ABEA_REFINE_MOM_START) return fail(ABEA_ERR_INVALID, "refine: unknown flags %d", flags);
  const bool mom = (flags & ABEA_REFINE_MOM_START) != 0;
  if (mom && !out->score) return fail(ABEA_ERR_INVALID, "refine: the start selection needs score");
  if (ms_per_step) *ms_per_step = 0.0;
  if (p->n == 0 || steps == 0) return ABEA_OK;
  HIP_TRY(hipSetDevice(p->ctx->device));
  abea_status_t st = ensure_events(p);
  if (st) return st;
  hipStream_t s = p->ctx->stream;
  // the batch's own scalings are updated in place; the next forward reads them
  abea_scalings_t *scal = const_cast<abea_scalings_t *>(p->batch.scalings);
  if (mom) {
    if ((st = p->rstart.ensure(sizeof(abea_scalings_t) * (size_t)p->n)) ||
        (st = p->rstate.ensure(sizeof(abea::RefineState) * (size_t)p->n)) ||
        (st = p->rstep.ensure(sizeof(int32_t))))
      return st;
  }
  const KernelArgs a = make_args(p, out);
  hipError_t e = hipEventRecord(p->evt[4], s);
  if (mom) {
    // the alternative start, and the step counter at 0 (a kernel, not a
    // memset, ahead of the captured steps)
    if (e == hipSuccess && (st = abea_plan_estimate_scalings(p, p->rstart.as<abea_scalings_t>())))
      return st;
    if (e == hipSuccess) e = abea::launch_refine_step_counter(p->rstep.as<int32_t>(), 0, s);
  }
  if (e != hipSuccess) return fail(ABEA_ERR_HIP, "refine: %s", hipGetErrorString(e));
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  abea_status_t cst = enqueue_forward(p, a, false);
  hipError_t le = hipSuccess;
  if (!cst)
    le = abea::launch_refine(a, scal, damping, mom ? p->rstep.as<int32_t>() : nullptr,
                             mom ? p->rstart.as<abea_scalings_t>() : nullptr,
                             mom ? p->rstate.as<abea::RefineState>() : nullptr, s);
  if (!cst && le == hipSuccess && mom)
    le = abea::launch_refine_step_counter(p->rstep.as<int32_t>(), -1, s);
  hipError_t ce = hipStreamEndCapture(s, &g);
  if (cst || le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return cst ? cst
               : fail(ABEA_ERR_HIP, "refine: graph capture failed: %s",
                      hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return fail(ABEA_ERR_HIP, "refine: hipGraphInstantiate failed: %s", hipGetErrorString(ie));
  const int launches = steps + (mom ? 1 : 0);
  for (int i = 0; i < launches && e == hipSuccess; ++i) e = hipGraphLaunch(ge, s);
  if (e == hipSuccess) e = hipEventRecord(p->evt[3], s);
  if (e == hipSuccess) e = hipEventSynchronize(p->evt[3]);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, p->evt[4], p->evt[3]);
  (void)hipGraphExecDestroy(ge);
  if (e != hipSuccess) return fail(ABEA_ERR_HIP, "refine: %s", hipGetErrorString(e));
  if (ms_per_step) *ms_per_step = ms / steps;
  p->have_times = false;
  return ABEA_OK;
}

abea_status_t abea_plan_get_tuning(const abea_plan_t *p, abea_tuning_t *t) {
  if (!p || !t) return fail(ABEA_ERR_INVALID, "get_tuning: null argument");
  *t = p->tune;
  return ABEA_OK;
}

abea_status_t abea_plan_set_tuning(abea_plan_t *p, const abea_tuning_t *t) {
  if (!p || !t) return fail(ABEA_ERR_INVALID, "set_tuning: null argument");
  const int w = t->core_waves_per_block;
  if ((w != 1 && w != 2 && w != 4) || t->core_waves_per_cu < 1 || t->core_waves_per_cu > 64 ||
      t->traceback_overlap < 32 || t->traceback_overlap > 1024 ||
      (t->fuse_traceback != 0 && t->fuse_traceback != 1))
    return fail(ABEA_ERR_INVALID, "set_tuning: unsupported values (%d, %d, %d, %d)", w,
                t->core_waves_per_cu, t->traceback_overlap, t->fuse_traceback);
  p->tune = *t;
  return ABEA_OK;
}

abea_status_t abea_plan_autotune(abea_plan_t *p, const abea_output_t *out, int reps,
                                 double *best_ms) {
  // The project autotuner: times candidate launch shapes on this plan's own
  // batch and keeps the fastest.  Axes: waves per workgroup and resident
  // waves per CU of the persistent core; whether each core wave traces its
  // own read back or a separate post kernel does; how far above its segment
  // each traceback lane starts.  Every candidate computes the same result;
  // only time differs.
  if (!p) return fail(ABEA_ERR_INVALID, "autotune: null plan");
  if (p->form != ABEA_PLAN_BATCHED) return fail(ABEA_ERR_INVALID, "autotune: batched plans only");
  abea_status_t st = check_out(p, out);
  if (st) return st;
  if (reps < 1) reps = 1;
  const abea_tuning_t saved = p->tune;
  // the fastest of n forwards, after one untimed forward at the new shape
  // (the first forward after a change of shape runs slow: without it, one
  // timed forward per candidate picked a shape 4% slower than the best)
  auto time_forward = [&](const abea_tuning_t &t, int n, double *ms) -> abea_status_t {
    p->tune = t;
    abea_status_t s1 = abea_plan_forward(p, out);
    if (s1) return s1;
    double best_run = 1e30;
    for (int r = 0; r < n; ++r) {
      abea_status_t s2 = abea_plan_forward(p, out);
      if (s2) return s2;
      double tm[4];
      if ((s2 = abea_plan_last_times(p, tm))) return s2;
      best_run = std::min(best_run, tm[0] + tm[1] + tm[2]);
    }
    *ms = best_run;
    return ABEA_OK;
  };
  struct Cand { int wpb, wpc, fuse, overlap; };
  // 20 waves per CU is the core's residency (five per SIMD); 32 queues the
  // rest behind them
  const Cand cands[] = {{1, 16, 0, 128}, {1, 24, 0, 128}, {1, 32, 0, 128}, {2, 32, 0, 128},
                        {4, 32, 0, 128}, {1, 16, 1, 128}, {1, 20, 1, 128}, {1, 24, 1, 128},
                        {1, 32, 1, 128}, {2, 32, 1, 128}, {4, 32, 1, 128}, {1, 24, 0, 64},
                        {1, 24, 0, 256}, {1, 24, 1, 64},  {1, 24, 1, 256}};
  // every candidate `reps` times, then the three fastest `reps` + 3 times more
  struct Timed {
    abea_tuning_t t;
    double ms;
  };
  std::vector<Timed> timed;
  for (const Cand &c : cands) {
    abea_tuning_t t = saved;
    t.use_graph = 0;
    t.core_waves_per_block = c.wpb;
    t.core_waves_per_cu = c.wpc;
    t.fuse_traceback = c.fuse;
    t.traceback_overlap = c.overlap;
    double ms = 0;
    if ((st = time_forward(t, reps, &ms))) {
      p->tune = saved;
      return st;
    }
    timed.push_back({t, ms});
  }
  std::stable_sort(timed.begin(), timed.end(),
                   [](const Timed &x, const Timed &y) { return x.ms < y.ms; });
  for (size_t i = 0; i < std::min<size_t>(3, timed.size()); ++i) {
    double ms = 0;
    if ((st = time_forward(timed[i].t, reps + 3, &ms))) {
      p->tune = saved;
      return st;
    }
    timed[i].ms = std::min(timed[i].ms, ms);
  }
  double best = 1e30;
  abea_tuning_t best_t = saved;
  for (size_t i = 0; i < std::min<size_t>(3, timed.size()); ++i)
    if (timed[i].ms < best) {
      best = timed[i].ms;
      best_t = timed[i].t;
    }
  best_t.use_graph = saved.use_graph;
  p->tune = best_t;
  if (best_ms) *best_ms = best;
  return ABEA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// f5c drop-in entries over host buffers.
namespace {

struct HostRun {
  abea_ctx *ctx = nullptr;
  abea_plan *plan = nullptr;
  DevBuf seq, seq_ptr, read_len, events, ev_ptr, n_events, scal, model, pairs, from, n_pairs,
      path_len, score;
  // align()/align_cuda() drop-ins: pinned inputs (bases, event means) and
  // outputs (pair counts, compacted pairs), and the compaction's offsets
  PinBuf h_seq, h_mean, h_np, h_pairs;
  DevBuf mean, coff, cpairs;
  DevBuf mom_order;  // MOM's read order (longest first)
  // post-align and MOM entries
  DevBuf pmeta, rec, bmap, n_rec, epb, flag, status, scal_out, terms, pstat, items;
  // getevents entries
  DevBuf ge_raw, ge_ev, ge_nev;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // drop-in phase timing
  ~HostRun() {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    delete plan;
    abea_ctx_destroy(ctx);
  }
};

// Host-buffer entry points share a pool of contexts (stream, plan, device
// buffers): a call takes one for its device and returns it when it ends, so
// the contexts held are bounded by the peak number of concurrent calls, not
// by the number of threads that ever called (f5c starts fresh worker threads
// for every batch).  The pool lives until the process ends.
std::mutex g_pool_mu;
std::vector<HostRun *> *g_pool = new std::vector<HostRun *>();
int32_t g_pool_created = 0;

struct HostLease {
  HostRun *r = nullptr;
  HostLease() = default;
  HostLease(const HostLease &) = delete;
  HostLease &operator=(const HostLease &) = delete;
  ~HostLease() {
    if (!r) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool->push_back(r);
  }
  abea_status_t acquire() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
      std::lock_guard<std::mutex> lk(g_pool_mu);
      for (size_t i = g_pool->size(); i-- > 0;)
        if ((*g_pool)[i]->ctx->device == dev) {
          r = (*g_pool)[i];
          g_pool->erase(g_pool->begin() + (std::ptrdiff_t)i);
          return ABEA_OK;
        }
    }
    HostRun *h = new HostRun();
    abea_status_t st = abea_ctx_create(&h->ctx, dev, nullptr);
    if (st) {
      delete h;
      return st;
    }
    h->plan = new abea_plan();
    h->plan->ctx = h->ctx;
    for (auto &e : h->ev)
      if (hipEventCreate(&e) != hipSuccess) {
        delete h;
        return fail(ABEA_ERR_HIP, "hipEventCreate failed");
      }
    r = h;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    ++g_pool_created;
    return ABEA_OK;
  }
};

// Phases of the calling thread's last abea_align / abea_align_db call (ms):
// flatten, host-to-device copies of the inputs, planning (per-read metadata
// and its upload), kernels, device-to-host copies, unflatten, whole call.
// The copies, planning and kernels are hipEvent-timed on the call's stream.
thread_local double t_host_ms[7] = {0, 0, 0, 0, 0, 0, 0};

double ms_since(const std::chrono::steady_clock::time_point &t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <class T>
abea_status_t upload(DevBuf &d, const T *h, size_t count, hipStream_t s) {
  abea_status_t st = d.ensure(sizeof(T) * std::max<size_t>(count, 1));
  if (st) return st;
  if (count) HIP_TRY(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
  return ABEA_OK;
}

// Runs fn(b, e) over [0, n) in up to 16 threads, split into contiguous
// ranges of about equal weight (wpre[i]: the weight of items before i;
// wpre[n] the total).  Small batches, and any failure to start a thread,
// run on the calling thread.
template <class F>
void parallel_ranges(int32_t n, const std::vector<int64_t> &wpre, F fn) {
  const int64_t total = wpre[(size_t)n];
  int T = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
  if (n < 2 || T < 2 || total < ((int64_t)1 << 22)) {
    fn(0, n);
    return;
  }
  T = std::min<int>(T, n);
  std::vector<std::thread> th;
  int32_t b = 0;
  try {
    for (int t = 0; t < T && b < n; ++t) {
      int32_t e = n;
      if (t + 1 < T) {
        const int64_t target = total * (t + 1) / T;
        e = (int32_t)(std::lower_bound(wpre.begin(), wpre.begin() + n, target) - wpre.begin());
        e = std::max(e, b + 1);
      }
      th.emplace_back(fn, b, e);
      b = e;
    }
  } catch (...) {
  }
  if (b < n) fn(b, n);  // what no thread took
  for (auto &x : th) x.join();
}

// A host-resident batch, flattened into the context's pinned staging: bases
// (h_seq, CSR by seq_ptr) and event means (h_mean, CSR by ev_ptr).
struct HostBatch {
  int32_t n = 0;
  std::vector<int64_t> seq_ptr, ev_ptr;
  std::vector<int32_t> read_len, n_events;
  int64_t n_bases = 0, n_ev = 0;
  bool staged = false;  // bases and event means already copied to r->seq / r->mean
};

// Runs a flattened batch.  On return the reads' pairs sit back to back in the
// pinned r->h_pairs, read i's at coff[i], with np[i] of them.
abea_status_t run_host_batch(HostRun *r, int form, const HostBatch &hb,
                             const abea_scalings_t *scal, const abea_model_t *model,
                             std::vector<int32_t> &np, std::vector<int64_t> &coff) {
  const int32_t n = hb.n;
  HIP_TRY(hipSetDevice(r->ctx->device));
  hipStream_t s = r->ctx->stream;
  abea_status_t st;
  HIP_TRY(hipEventRecord(r->ev[0], s));
  if ((!hb.staged && (st = upload(r->seq, r->h_seq.as<char>(), (size_t)hb.n_bases, s))) ||
      (st = upload(r->seq_ptr, hb.seq_ptr.data(), hb.seq_ptr.size(), s)) ||
      (st = upload(r->read_len, hb.read_len.data(), hb.read_len.size(), s)) ||
      (!hb.staged && (st = upload(r->mean, r->h_mean.as<float>(), (size_t)hb.n_ev, s))) ||
      (st = upload(r->ev_ptr, hb.ev_ptr.data(), hb.ev_ptr.size(), s)) ||
      (st = upload(r->n_events, hb.n_events.data(), hb.n_events.size(), s)) ||
      (st = upload(r->scal, scal, (size_t)n, s)) ||
      (st = upload(r->model, model, (size_t)ABEA_NUM_KMER, s)))
    return st;
  HIP_TRY(hipEventRecord(r->ev[1], s));
  abea_batch_t b{};
  b.n_reads = n;
  b.seq = r->seq.as<char>();
  b.seq_ptr = r->seq_ptr.as<int64_t>();
  b.read_len = r->read_len.as<int32_t>();
  b.events = nullptr;
  b.ev_mean = r->mean.as<float>();
  b.ev_ptr = r->ev_ptr.as<int64_t>();
  b.n_events = r->n_events.as<int32_t>();
  b.scalings = r->scal.as<abea_scalings_t>();
  b.model = r->model.as<abea_model_t>();
  b.read_len_host = hb.read_len.data();
  b.n_events_host = hb.n_events.data();
  b.seq_ptr_host = hb.seq_ptr.data();
  b.ev_ptr_host = hb.ev_ptr.data();
  if ((st = plan_bind(r->plan, &b, form))) return st;
  const int64_t cap = r->plan->total_pairs;
  if ((st = r->pairs.ensure(sizeof(abea_aligned_pair_t) * (size_t)std::max<int64_t>(cap, 1)))) return st;
  if ((st = r->n_pairs.ensure(sizeof(int32_t) * (size_t)std::max(n, 1)))) return st;
  if ((st = r->h_np.ensure(sizeof(int32_t) * (size_t)std::max(n, 1)))) return st;
  abea_output_t out{};
  out.pairs = r->pairs.as<abea_aligned_pair_t>();
  out.n_pairs = r->n_pairs.as<int32_t>();
  HIP_TRY(hipEventRecord(r->ev[2], s));
  if ((st = abea_plan_forward(r->plan, &out))) return st;
  HIP_TRY(hipEventRecord(r->ev[3], s));
  // the counts, then only the pairs (compacted on the device)
  HIP_TRY(hipMemcpyAsync(r->h_np.p, out.n_pairs, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  np.assign(r->h_np.as<int32_t>(), r->h_np.as<int32_t>() + n);
  coff.resize((size_t)n + 1);
  int64_t tot = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int64_t slots = (int64_t)hb.n_events[i] + std::max<int64_t>((int64_t)hb.read_len[i] - 5, 0);
    if (np[i] < 0 || np[i] > slots)
      return fail(ABEA_ERR_HIP, "align: read %d returned %d pairs for %lld slots", i, np[i],
                  (long long)slots);
    coff[i] = tot;
    tot += np[i];
  }
  coff[n] = tot;
  if (tot) {
    if ((st = upload(r->coff, coff.data(), (size_t)n, s))) return st;
    if ((st = r->cpairs.ensure(sizeof(abea_aligned_pair_t) * (size_t)tot))) return st;
    if ((st = r->h_pairs.ensure(sizeof(abea_aligned_pair_t) * (size_t)tot))) return st;
    HIP_TRY(abea::launch_compact_pairs(r->plan->meta.as<ReadMeta>(), out.pairs, out.n_pairs,
                                       r->coff.as<int64_t>(), r->cpairs.as<abea_aligned_pair_t>(),
                                       n, s));
    HIP_TRY(hipMemcpyAsync(r->h_pairs.p, r->cpairs.p, sizeof(abea_aligned_pair_t) * (size_t)tot,
                           hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipEventRecord(r->ev[4], s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  for (int i = 0; i < 4; ++i) {
    HIP_TRY(hipEventElapsedTime(&ms, r->ev[i], r->ev[i + 1]));
    t_host_ms[1 + i] = ms;
  }
  return ABEA_OK;
}

// Concurrent abea_align() calls on one device are gathered into batches.
// f5c calls align() from its worker threads, one read per call
// (GPU/f5c.h:57-59), and one read is one wave walking its bands, so calls run
// one by one leave the GPU nearly idle.  A call queues its read; if no batch
// is running on its device, the call takes every queued read that shares the
// oldest one's model, runs them as one batched forward (whose results equal
// the per-read form's), hands each caller its pairs and wakes them.  Calls
// that arrive meanwhile wait for the next batch, which one of them runs.  A
// call alone runs a batch of one.
struct AlignReq {
  const char *seq = nullptr;
  int32_t len = 0;
  const abea_event_t *ev = nullptr;
  int32_t n_e = 0;
  abea_scalings_t sc{};
  const abea_model_t *model = nullptr;
  abea_aligned_pair_t *out = nullptr;
  int32_t result = -1;
  bool done = false;
  std::string err;
  double times[7] = {0, 0, 0, 0, 0, 0, 0};  // the batch's phases
};

struct Collector {
  std::mutex mu;
  std::condition_variable cv;
  std::vector<AlignReq *> pending;
  bool busy = false;  // a batch is running on this device
  int callers = 0;    // calls inside abea_align on this device
  int seen = 0;       // most callers at once since the last batch started
  int target = 1;     // the same over the interval before (a batch's target size)
};

// How long a call about to run a batch waits for more calls to queue, when
// fewer are queued than ran at once before: the callers of the last batch
// come back a few microseconds after it ends, and a batch of one read leaves
// the GPU nearly idle for a whole walk (milliseconds).
constexpr auto kGatherWait = std::chrono::microseconds(200);

constexpr int kMaxHostDevices = 64;
Collector g_coll[kMaxHostDevices];

// Runs one batch of queued reads; sets every request's result (pairs, or -1
// and the error).  Catches everything: the waiting callers depend on it.
void run_align_batch(const std::vector<AlignReq *> &q) noexcept {
  const auto t0 = std::chrono::steady_clock::now();
  const int32_t n = (int32_t)q.size();
  abea_status_t st = ABEA_OK;
  try {
    HostLease lease;
    st = lease.acquire();
    HostRun *r = lease.r;
    HostBatch hb;
    std::vector<abea_scalings_t> scal(n);
    if (!st) {
      hb.n = n;
      hb.seq_ptr.resize(n);
      hb.ev_ptr.resize(n);
      hb.read_len.resize(n);
      hb.n_events.resize(n);
      for (int32_t i = 0; i < n; ++i) {
        hb.seq_ptr[i] = hb.n_bases;
        hb.ev_ptr[i] = hb.n_ev;
        hb.read_len[i] = q[i]->len;
        hb.n_events[i] = q[i]->n_e;
        scal[i] = q[i]->sc;
        hb.n_bases += std::max(q[i]->len, 0);
        hb.n_ev += q[i]->n_e;
      }
      if (!(st = r->h_seq.ensure((size_t)std::max<int64_t>(hb.n_bases, 1))))
        st = r->h_mean.ensure(sizeof(float) * (size_t)std::max<int64_t>(hb.n_ev, 1));
    }
    std::vector<int32_t> np;
    std::vector<int64_t> coff;
    if (!st) {
      char *hs = r->h_seq.as<char>();
      float *hm = r->h_mean.as<float>();
      for (int32_t i = 0; i < n; ++i) {
        if (q[i]->len > 0) std::memcpy(hs + hb.seq_ptr[i], q[i]->seq, (size_t)q[i]->len);
        float *dst = hm + hb.ev_ptr[i];
        for (int32_t j = 0; j < q[i]->n_e; ++j) dst[j] = q[i]->ev[j].mean;
      }
      t_host_ms[0] = ms_since(t0);
      st = run_host_batch(r, ABEA_PLAN_BATCHED, hb, scal.data(), q[0]->model, np, coff);
    }
    if (!st) {
      const auto t4 = std::chrono::steady_clock::now();
      const abea_aligned_pair_t *hp = r->h_pairs.as<abea_aligned_pair_t>();
      for (int32_t i = 0; i < n; ++i) {
        int32_t k = np[i];
        if ((int64_t)k > 2 * (int64_t)q[i]->n_e) k = 0;  // would overrun f5c's 2*n_events buffer
        if (k) std::memcpy(q[i]->out, hp + coff[i], sizeof(abea_aligned_pair_t) * (size_t)k);
        q[i]->result = k;
      }
      t_host_ms[5] = ms_since(t4);
      t_host_ms[6] = ms_since(t0);
      for (AlignReq *x : q) std::copy(t_host_ms, t_host_ms + 7, x->times);
      return;
    }
  } catch (const std::exception &e) {
    fail(ABEA_ERR_NOMEM, "abea_align: %s", e.what());
  }
  for (AlignReq *x : q) {
    x->result = -1;
    x->err = g_err;
  }
}

}  // namespace

extern "C" {

int32_t abea_align(abea_aligned_pair_t *out_2, char *sequence, int32_t sequence_len,
                   abea_event_table events, abea_model_t *models, abea_scalings_t scaling,
                   float sample_rate) {
  (void)sample_rate;  // unused by the reference too, GPU/align_cuda.cl:84-88
  if (!out_2 || !sequence || !models || (events.n && !events.event)) {
    fail(ABEA_ERR_INVALID, "abea_align: null argument");
    return -1;
  }
  if (events.n > (size_t)0x7fffffff) {
    fail(ABEA_ERR_INVALID, "abea_align: %zu events (need < 2^31)", events.n);
    return -1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxHostDevices) {
    (void)hipGetLastError();
    fail(ABEA_ERR_HIP, "abea_align: no usable current device (%d)", dev);
    return -1;
  }
  AlignReq req;
  req.seq = sequence;
  req.len = sequence_len;
  req.ev = events.event;
  req.n_e = (int32_t)events.n;
  req.sc = scaling;
  req.model = models;
  req.out = out_2;
  Collector &c = g_coll[dev];
  try {
    std::unique_lock<std::mutex> lk(c.mu);
    c.pending.push_back(&req);
    c.seen = std::max(c.seen, ++c.callers);
    c.cv.notify_all();  // a call waiting to gather may now have enough
    while (!req.done) {
      if (c.busy) {
        c.cv.wait(lk);
        continue;
      }
      if ((int)c.pending.size() < c.target) {
        const auto until = std::chrono::steady_clock::now() + kGatherWait;
        while (!c.busy && !req.done && (int)c.pending.size() < c.target &&
               c.cv.wait_until(lk, until) != std::cv_status::timeout) {
        }
        if (c.busy || req.done) continue;
      }
      c.target = std::max(1, c.seen);
      c.seen = c.callers;
      // run every queued read with the oldest one's model as one batch
      std::vector<AlignReq *> batch, rest;
      batch.reserve(c.pending.size());
      rest.reserve(c.pending.size());
      const abea_model_t *mdl = c.pending.front()->model;
      for (AlignReq *q : c.pending) (q->model == mdl ? batch : rest).push_back(q);
      c.pending.swap(rest);
      c.busy = true;
      lk.unlock();
      run_align_batch(batch);
      lk.lock();
      for (AlignReq *q : batch) q->done = true;
      c.busy = false;
      c.cv.notify_all();
    }
    --c.callers;
  } catch (const std::exception &e) {
    // Only the queues' allocations can throw, before this call's read left
    // the queue (the batch itself catches everything): take it out again.
    std::lock_guard<std::mutex> g(c.mu);
    auto it = std::find(c.pending.begin(), c.pending.end(), &req);
    if (it != c.pending.end()) {
      c.pending.erase(it);
      --c.callers;
    }
    fail(ABEA_ERR_NOMEM, "abea_align: %s", e.what());
    return -1;
  }
  std::copy(req.times, req.times + 7, t_host_ms);
  if (req.result < 0) abea::set_last_error(req.err.c_str());
  return req.result;
}

abea_status_t abea_align_db(int32_t n_bam_rec, char **read, const int32_t *read_len,
                            const abea_event_table *et, const abea_model_t *model,
                            const abea_scalings_t *scalings,
                            abea_aligned_pair_t **event_align_pairs,
                            int32_t *n_event_align_pairs) {
  if (n_bam_rec < 0 || (n_bam_rec > 0 && (!read || !read_len || !et || !model || !scalings ||
                                          !event_align_pairs || !n_event_align_pairs)))
    return fail(ABEA_ERR_INVALID, "abea_align_db: null argument");
  const auto t0 = std::chrono::steady_clock::now();
  const int32_t n = n_bam_rec;
  HostBatch hb;
  hb.n = n;
  hb.seq_ptr.resize(n);
  hb.ev_ptr.resize(n);
  hb.read_len.assign(read_len, read_len + n);
  hb.n_events.resize(n);
  std::vector<int64_t> wpre((size_t)n + 1);
  for (int32_t i = 0; i < n; ++i) {
    if ((read_len[i] > 0 && !read[i]) || (et[i].n && !et[i].event))
      return fail(ABEA_ERR_INVALID, "abea_align_db: read %d: null bases or events", i);
    if (et[i].n > (size_t)0x7fffffff)
      return fail(ABEA_ERR_INVALID, "abea_align_db: read %d: %zu events (need < 2^31)", i,
                  et[i].n);
    hb.seq_ptr[i] = hb.n_bases;
    hb.ev_ptr[i] = hb.n_ev;
    hb.n_events[i] = (int32_t)et[i].n;
    wpre[i] = 24 * hb.n_ev + hb.n_bases;  // bytes the flatten reads before read i
    hb.n_bases += std::max(read_len[i], 0);
    hb.n_ev += (int64_t)et[i].n;
  }
  wpre[n] = 24 * hb.n_ev + hb.n_bases;
  HostLease lease;
  abea_status_t st = lease.acquire();
  if (st) return st;
  HostRun *r = lease.r;
  if ((st = r->h_seq.ensure((size_t)std::max<int64_t>(hb.n_bases, 1))) ||
      (st = r->h_mean.ensure(sizeof(float) * (size_t)std::max<int64_t>(hb.n_ev, 1))))
    return st;
  // flatten, as GPU/align_cuda.c:432-492 does, into pinned staging: the bases
  // (no NUL terminators needed) and the event means, the one field of the
  // event records the alignment reads.  Each thread copies its reads to the
  // device in four parts as it goes, so the transfer runs under the flatten
  // instead of after it.
  HIP_TRY(hipSetDevice(r->ctx->device));
  if ((st = r->seq.ensure((size_t)std::max<int64_t>(hb.n_bases, 1))) ||
      (st = r->mean.ensure(sizeof(float) * (size_t)std::max<int64_t>(hb.n_ev, 1))))
    return st;
  char *hs = r->h_seq.as<char>();
  float *hm = r->h_mean.as<float>();
  char *ds = r->seq.as<char>();
  float *dm = r->mean.as<float>();
  const hipStream_t cs = r->ctx->stream;
  std::atomic<int> copy_err{0};
  const int dev = r->ctx->device;
  parallel_ranges(n, wpre, [&](int32_t b, int32_t e) {
    if (hipSetDevice(dev) != hipSuccess) copy_err = 1;  // this thread's current device
    int32_t i = b;
    for (int part = 1; part <= 4 && i < e; ++part) {
      int32_t j = e;
      if (part < 4) {
        const int64_t target = wpre[b] + (wpre[e] - wpre[b]) * part / 4;
        j = (int32_t)(std::lower_bound(wpre.begin() + i, wpre.begin() + e, target) -
                      wpre.begin());
        j = std::max(j, i + 1);
      }
      for (int32_t k = i; k < j; ++k) {
        if (read_len[k] > 0) std::memcpy(hs + hb.seq_ptr[k], read[k], (size_t)read_len[k]);
        const abea_event_t *src = et[k].event;
        float *dst = hm + hb.ev_ptr[k];
        for (size_t q = 0; q < et[k].n; ++q) dst[q] = src[q].mean;
      }
      const int64_t sb = hb.seq_ptr[i], se = j < n ? hb.seq_ptr[j] : hb.n_bases;
      const int64_t eb = hb.ev_ptr[i], ee = j < n ? hb.ev_ptr[j] : hb.n_ev;
      if (se > sb && hipMemcpyAsync(ds + sb, hs + sb, (size_t)(se - sb), hipMemcpyHostToDevice,
                                    cs) != hipSuccess)
        copy_err = 1;
      if (ee > eb && hipMemcpyAsync(dm + eb, hm + eb, sizeof(float) * (size_t)(ee - eb),
                                    hipMemcpyHostToDevice, cs) != hipSuccess)
        copy_err = 1;
      i = j;
    }
  });
  if (copy_err) {
    (void)hipGetLastError();
    return fail(ABEA_ERR_HIP, "abea_align_db: copy of the flattened inputs failed");
  }
  hb.staged = true;
  t_host_ms[0] = ms_since(t0);
  std::vector<int32_t> np;
  std::vector<int64_t> coff;
  if ((st = run_host_batch(r, ABEA_PLAN_BATCHED, hb, scalings, model, np, coff))) return st;
  const auto t4 = std::chrono::steady_clock::now();
  // unflatten into callee-malloc'd per-read arrays, GPU/align_cuda.c:926-931
  for (int32_t i = 0; i < n; ++i) {
    event_align_pairs[i] =
        (abea_aligned_pair_t *)malloc(sizeof(abea_aligned_pair_t) * (size_t)std::max(np[i], 1));
    if (!event_align_pairs[i]) {
      // the caller owns nothing from a failed call
      for (int32_t j = 0; j < i; ++j) free(event_align_pairs[j]);
      for (int32_t j = 0; j < n; ++j) {
        event_align_pairs[j] = nullptr;
        n_event_align_pairs[j] = 0;
      }
      return fail(ABEA_ERR_NOMEM, "abea_align_db: malloc failed");
    }
    n_event_align_pairs[i] = np[i];
  }
  const abea_aligned_pair_t *hp = r->h_pairs.as<abea_aligned_pair_t>();
  parallel_ranges(n, coff, [&](int32_t b, int32_t e) {
    for (int32_t i = b; i < e; ++i)
      if (np[i])
        std::memcpy(event_align_pairs[i], hp + coff[i], sizeof(abea_aligned_pair_t) * (size_t)np[i]);
  });
  t_host_ms[5] = ms_since(t4);
  t_host_ms[6] = ms_since(t0);
  return ABEA_OK;
}

void abea_host_last_times(double times_ms[7]) {
  if (!times_ms) return;
  for (int i = 0; i < 7; ++i) times_ms[i] = t_host_ms[i];
}

int32_t abea_host_contexts(void) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  return g_pool_created;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// f5c's steps either side of align() over host buffers: flatten, upload, one
// launch, read back.  Per-read metadata for these kernels is built here (no
// plan: they need no trace or staging workspace).
namespace {

struct HostPost {
  int32_t n = 0;
  std::vector<ReadMeta> meta;
  std::vector<char> seq;
  std::vector<int64_t> seq_ptr;
  std::vector<abea_event_t> ev;
};

// Reads i = 0..n-1 with read_len[i] bases and n_e[i] events; pair/record
// slots by pair_cnt[i] (may be null) and map slots by max(read_len - 5, 0).
void host_post_layout(HostPost &h, int32_t n, char *const *read, const int32_t *read_len,
                      const abea_event_table *et, const int32_t *pair_cnt) {
  h.n = n;
  h.meta.assign(n, ReadMeta{});
  h.seq_ptr.resize(n);
  int64_t sl = 0, se = 0, pr = 0, bm = 0;
  for (int32_t i = 0; i < n; ++i) {
    ReadMeta &m = h.meta[i];
    m.n_k = read_len[i] - ABEA_KMER_SIZE + 1;
    m.n_e = et ? (int32_t)et[i].n : 0x7fffffff;
    m.ev_off = se;
    m.pair_off = pr;
    m.bmap_off = bm;
    h.seq_ptr[i] = sl;
    sl += std::max(read_len[i], 0);
    se += et ? (int64_t)et[i].n : 0;
    pr += pair_cnt ? std::max(pair_cnt[i], 0) : 0;
    bm += std::max(m.n_k, 0);
  }
  h.seq.resize((size_t)sl);
  h.ev.resize((size_t)se);
  for (int32_t i = 0; i < n; ++i) {
    if (read_len[i] > 0) std::memcpy(h.seq.data() + h.seq_ptr[i], read[i], (size_t)read_len[i]);
    if (et && et[i].n)
      std::memcpy(h.ev.data() + h.meta[i].ev_off, et[i].event, sizeof(abea_event_t) * et[i].n);
  }
}

abea_status_t host_post_upload(HostRun *r, const HostPost &h, const abea_model_t *model,
                               hipStream_t s) {
  abea_status_t st;
  if ((st = upload(r->pmeta, h.meta.data(), h.meta.size(), s))) return st;
  if ((st = upload(r->seq, h.seq.data(), h.seq.size(), s))) return st;
  if ((st = upload(r->seq_ptr, h.seq_ptr.data(), h.seq_ptr.size(), s))) return st;
  if ((st = upload(r->events, h.ev.data(), h.ev.size(), s))) return st;
  if (model && (st = upload(r->model, model, (size_t)ABEA_NUM_KMER, s))) return st;
  return ABEA_OK;
}

template <class T>
abea_status_t download(T *h, const DevBuf &d, size_t count, hipStream_t s) {
  if (count) HIP_TRY(hipMemcpyAsync(h, d.p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
  return ABEA_OK;
}

// MOM over n reads into scal_h.
abea_status_t run_mom(int32_t n, char *const *read, const int32_t *read_len,
                      const abea_event_table *et, const abea_model_t *model,
                      abea_scalings_t *scal_h) {
  HostLease lease;
  abea_status_t st = lease.acquire();
  if (st) return st;
  HostRun *r = lease.r;
  HIP_TRY(hipSetDevice(r->ctx->device));
  hipStream_t s = r->ctx->stream;
  HostPost h;
  host_post_layout(h, n, read, read_len, et, nullptr);
  if ((st = host_post_upload(r, h, model, s))) return st;
  if ((st = r->scal_out.ensure(sizeof(abea_scalings_t) * (size_t)std::max(n, 1)))) return st;
  // longest serial chains first
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  auto chain = [&](int32_t i) {
    return std::max<int64_t>((int64_t)et[i].n, std::max<int64_t>((int64_t)read_len[i] - 5, 0));
  };
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return chain(x) > chain(y); });
  if ((st = upload(r->mom_order, order.data(), order.size(), s))) return st;
  abea::MomArgs a{};
  a.n_reads = n;
  a.order = r->mom_order.as<int32_t>();
  a.meta = r->pmeta.as<ReadMeta>();
  a.seq = r->seq.as<char>();
  a.seq_ptr = r->seq_ptr.as<int64_t>();
  a.events = r->events.as<abea_event_t>();
  a.model = r->model.as<abea_model_t>();
  a.out = r->scal_out.as<abea_scalings_t>();
  HIP_TRY(abea::launch_mom(a, s));
  if ((st = download(scal_h, r->scal_out, (size_t)n, s))) return st;
  HIP_TRY(hipStreamSynchronize(s));
  return ABEA_OK;
}

// postalign (+ recalibration when et and model are given) over n reads whose
// pairs are given per read.  Results stay in r's buffers; flag_h, n_rec_h,
// epb_h and scal_h (when recalibrating) are read back.
abea_status_t run_postalign(HostRun *r, HostPost &h, char *const *read, const int32_t *read_len,
                            const abea_event_table *et, const abea_model_t *model,
                            abea_aligned_pair_t *const *pairs, const int32_t *n_pairs,
                            abea_scalings_t *scal_h, int recalibrate,
                            std::vector<int32_t> &flag_h, std::vector<int32_t> &n_rec_h,
                            std::vector<double> &epb_h) {
  const int32_t n = h.n;
  hipStream_t s = r->ctx->stream;
  abea_status_t st;
  if ((st = host_post_upload(r, h, recalibrate ? model : nullptr, s))) return st;
  int64_t np_tot = 0, bm_tot = 0;
  for (int32_t i = 0; i < n; ++i) {
    np_tot += std::max(n_pairs[i], 0);
    bm_tot += std::max(h.meta[i].n_k, 0);
  }
  std::vector<abea_aligned_pair_t> flat((size_t)np_tot);
  for (int32_t i = 0; i < n; ++i)
    if (n_pairs[i] > 0)
      std::memcpy(flat.data() + h.meta[i].pair_off, pairs[i],
                  sizeof(abea_aligned_pair_t) * (size_t)n_pairs[i]);
  if ((st = upload(r->pairs, flat.data(), flat.size(), s))) return st;
  if ((st = upload(r->n_pairs, n_pairs, (size_t)n, s))) return st;
  if (recalibrate && (st = upload(r->scal, scal_h, (size_t)n, s))) return st;
  if (!recalibrate && (st = r->scal.ensure(sizeof(abea_scalings_t) * (size_t)std::max(n, 1))))
    return st;
  if ((st = r->rec.ensure(sizeof(abea_event_alignment_t) * (size_t)std::max<int64_t>(np_tot, 1))))
    return st;
  if ((st = r->bmap.ensure(sizeof(abea_index_pair_t) * (size_t)std::max<int64_t>(bm_tot, 1))))
    return st;
  if ((st = r->n_rec.ensure(sizeof(int32_t) * (size_t)std::max(n, 1)))) return st;
  if ((st = r->epb.ensure(sizeof(double) * (size_t)std::max(n, 1)))) return st;
  if ((st = r->flag.ensure(sizeof(int32_t) * (size_t)std::max(n, 1)))) return st;
  if ((st = r->scal_out.ensure(sizeof(abea_scalings_t) * (size_t)std::max(n, 1)))) return st;
  if (recalibrate &&
      (st = r->terms.ensure(sizeof(float4) * (size_t)std::max<int64_t>(bm_tot, 1))))
    return st;
  std::vector<int2> items;
  const int seg = abea::postalign_segment_pairs();
  for (int32_t i = 0; i < n; ++i)
    for (int64_t q = 0; q * seg < (int64_t)n_pairs[i]; ++q) items.push_back(make_int2(i, (int)q));
  if ((st = upload(r->items, items.data(), items.size(), s))) return st;
  if ((st = r->pstat.ensure(sizeof(int32_t) * (size_t)std::max(n, 1)))) return st;
  HIP_TRY(hipMemsetAsync(r->pstat.p, 0, sizeof(int32_t) * (size_t)std::max(n, 1), s));
  abea::PostArgs a{};
  a.n_reads = n;
  a.meta = r->pmeta.as<ReadMeta>();
  a.seq = r->seq.as<char>();
  a.seq_ptr = r->seq_ptr.as<int64_t>();
  a.events = r->events.as<abea_event_t>();
  a.model = r->model.as<abea_model_t>();
  a.scal_in = r->scal.as<abea_scalings_t>();
  a.pairs = r->pairs.as<abea_aligned_pair_t>();
  a.n_pairs = r->n_pairs.as<int32_t>();
  a.rec = r->rec.as<abea_event_alignment_t>();
  a.bmap = r->bmap.as<abea_index_pair_t>();
  a.n_rec = r->n_rec.as<int32_t>();
  a.epb = r->epb.as<double>();
  a.scal_out = r->scal_out.as<abea_scalings_t>();
  a.flag = r->flag.as<int32_t>();
  a.items = r->items.as<int2>();
  a.n_items = (int32_t)items.size();
  a.pstat = r->pstat.as<int32_t>();
  a.terms = recalibrate ? r->terms.as<float4>() : nullptr;
  a.recalibrate = recalibrate;
  a.scale_var = 1;
  HIP_TRY(abea::launch_postalign(a, s));
  flag_h.resize(n);
  n_rec_h.resize(n);
  epb_h.resize(n);
  if ((st = download(flag_h.data(), r->flag, (size_t)n, s))) return st;
  if ((st = download(n_rec_h.data(), r->n_rec, (size_t)n, s))) return st;
  if ((st = download(epb_h.data(), r->epb, (size_t)n, s))) return st;
  if (recalibrate && (st = download(scal_h, r->scal_out, (size_t)n, s))) return st;
  HIP_TRY(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n; ++i)
    if (flag_h[i] & abea::kBadPath)
      return fail(ABEA_ERR_INVALID,
                  "postalign: the pairs of read %d are not an alignment path (steps to the "
                  "next event and/or k-mer, k-mers in [0, %d)%s)",
                  i, h.meta[i].n_k, recalibrate ? ", events in range" : "");
  return ABEA_OK;
}

}  // namespace

extern "C" {

abea_scalings_t abea_estimate_scalings_using_mom(char *sequence, int32_t sequence_len,
                                                 abea_model_t *pore_model, abea_event_table et) {
  abea_scalings_t out;
  out.scale = out.shift = out.var = out.log_var = std::nanf("");
  if (!sequence || !pore_model || (et.n && !et.event) || sequence_len < 0) {
    fail(ABEA_ERR_INVALID, "estimate_scalings_using_mom: null argument");
    return out;
  }
  char *rd[1] = {sequence};
  if (run_mom(1, rd, &sequence_len, &et, pore_model, &out)) out.scale = out.shift = std::nanf("");
  return out;
}

abea_status_t abea_getevents_db(int32_t n_reads, const size_t *nsample, float *const *rawptr,
                                abea_event_table *et) {
  if (n_reads < 0 || (n_reads > 0 && (!nsample || !rawptr || !et)))
    return fail(ABEA_ERR_INVALID, "getevents_db: null argument or negative read count");
  for (int32_t i = 0; i < n_reads; ++i) {
    et[i] = abea_event_table{0, 0, 0, nullptr};
    if (nsample[i] && !rawptr[i]) return fail(ABEA_ERR_INVALID, "getevents: read %d: null rawptr", i);
    if (nsample[i] >= ((size_t)1 << 31))
      return fail(ABEA_ERR_INVALID, "getevents: read %d has %zu samples (need < 2^31)", i,
                  nsample[i]);
  }
  if (n_reads == 0) return ABEA_OK;
  HostLease lease;
  abea_status_t st = lease.acquire();
  if (st) return st;
  HostRun *r = lease.r;
  HIP_TRY(hipSetDevice(r->ctx->device));
  hipStream_t s = r->ctx->stream;
  std::vector<int64_t> off((size_t)n_reads + 1, 0);
  for (int32_t i = 0; i < n_reads; ++i) off[i + 1] = off[i] + (int64_t)nsample[i];
  const int64_t total = off[n_reads];
  if ((st = r->ge_raw.ensure(sizeof(float) * (size_t)std::max<int64_t>(total, 1))) ||
      (st = r->ge_ev.ensure(sizeof(abea_event_t) * (size_t)(total + n_reads))) ||
      (st = r->ge_nev.ensure(sizeof(int64_t) * (size_t)n_reads)))
    return st;
  for (int32_t i = 0; i < n_reads; ++i)
    if (nsample[i])
      HIP_TRY(hipMemcpyAsync(r->ge_raw.as<float>() + off[i], rawptr[i], sizeof(float) * nsample[i],
                             hipMemcpyHostToDevice, s));
  if ((st = ge_enqueue(r->ctx, n_reads, off.data(), r->ge_raw.as<float>(),
                       r->ge_ev.as<abea_event_t>(), r->ge_nev.as<int64_t>())))
    return st;
  std::vector<int64_t> cnt((size_t)n_reads);
  HIP_TRY(hipMemcpyAsync(cnt.data(), r->ge_nev.p, sizeof(int64_t) * (size_t)n_reads,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n_reads; ++i) {
    if (cnt[i] <= 0) continue;
    abea_event_t *e = (abea_event_t *)malloc(sizeof(abea_event_t) * (size_t)cnt[i]);
    if (!e) {
      for (int32_t j = 0; j < i; ++j) {
        free(et[j].event);
        et[j] = abea_event_table{0, 0, 0, nullptr};
      }
      return fail(ABEA_ERR_NOMEM, "getevents: malloc of %lld events failed", (long long)cnt[i]);
    }
    HIP_TRY(hipMemcpyAsync(e, r->ge_ev.as<abea_event_t>() + off[i] + i,
                           sizeof(abea_event_t) * (size_t)cnt[i], hipMemcpyDeviceToHost, s));
    et[i] = abea_event_table{(size_t)cnt[i], 0, (size_t)cnt[i], e};
  }
  HIP_TRY(hipStreamSynchronize(s));
  return ABEA_OK;
}

abea_event_table abea_getevents(size_t nsample, float *rawptr) {
  abea_event_table et{0, 0, 0, nullptr};
  if (nsample && !rawptr) {
    fail(ABEA_ERR_INVALID, "getevents: null rawptr");
    return et;
  }
  if (abea_getevents_db(1, &nsample, &rawptr, &et)) et = abea_event_table{0, 0, 0, nullptr};
  return et;
}

abea_status_t abea_estimate_scalings_db(int32_t n_bam_rec, char **read, const int32_t *read_len,
                                        const abea_event_table *et, const abea_model_t *model,
                                        abea_scalings_t *scalings) {
  if (n_bam_rec < 0 || (n_bam_rec > 0 && (!read || !read_len || !et || !model || !scalings)))
    return fail(ABEA_ERR_INVALID, "estimate_scalings_db: null argument");
  if (n_bam_rec == 0) return ABEA_OK;
  return run_mom(n_bam_rec, read, read_len, et, model, scalings);
}

int32_t abea_postalign(abea_event_alignment_t *alignment, abea_index_pair_t *base_to_event_map,
                       double *events_per_base, char *sequence, int32_t n_kmers,
                       abea_aligned_pair_t *event_alignment, int32_t n_events) {
  if (!alignment || !base_to_event_map || !events_per_base || !sequence || !event_alignment) {
    fail(ABEA_ERR_INVALID, "postalign: null argument");
    return -1;
  }
  if (n_kmers <= 0 || n_events <= 0) {
    fail(ABEA_ERR_INVALID, "postalign: needs n_kmers > 0 and n_events > 0 pairs (got %d, %d)",
         n_kmers, n_events);
    return -1;
  }
  HostLease lease;
  if (lease.acquire()) return -1;
  HostRun *r = lease.r;
  if (hipSetDevice(r->ctx->device) != hipSuccess) {
    fail(ABEA_ERR_HIP, "postalign: hipSetDevice failed");
    return -1;
  }
  const int32_t len = n_kmers + ABEA_KMER_SIZE - 1;
  char *rd[1] = {sequence};
  HostPost h;
  host_post_layout(h, 1, rd, &len, nullptr, &n_events);
  std::vector<int32_t> flag, n_rec;
  std::vector<double> epb;
  abea_aligned_pair_t *pp[1] = {event_alignment};
  if (run_postalign(r, h, rd, &len, nullptr, nullptr, pp, &n_events, nullptr, 0, flag, n_rec,
                    epb))
    return -1;
  hipStream_t s = r->ctx->stream;
  if (download(alignment, r->rec, (size_t)n_rec[0], s) ||
      download(base_to_event_map, r->bmap, (size_t)n_kmers, s))
    return -1;
  if (hipStreamSynchronize(s) != hipSuccess) {
    fail(ABEA_ERR_HIP, "postalign: stream synchronize failed");
    return -1;
  }
  *events_per_base = epb[0];
  return n_rec[0];
}

bool abea_recalibrate_model(abea_model_t *pore_model, abea_event_table et,
                            abea_scalings_t *scallings,
                            const abea_event_alignment_t *alignment_output,
                            int32_t num_alignments, bool scale_var) {
  if (!pore_model || !scallings || (num_alignments > 0 && !alignment_output) ||
      (et.n && !et.event) || num_alignments < 0) {
    fail(ABEA_ERR_INVALID, "recalibrate_model: null argument");
    return false;
  }
  HostLease lease;
  if (lease.acquire()) return false;
  HostRun *r = lease.r;
  hipStream_t s = r->ctx->stream;
  auto go = [&]() -> abea_status_t {
    HIP_TRY(hipSetDevice(r->ctx->device));
    abea_status_t st;
    if ((st = upload(r->model, pore_model, (size_t)ABEA_NUM_KMER, s))) return st;
    if ((st = upload(r->events, et.event, (size_t)et.n, s))) return st;
    if ((st = upload(r->rec, alignment_output, (size_t)num_alignments, s))) return st;
    if ((st = upload(r->scal, scallings, 1, s))) return st;
    if ((st = r->status.ensure(sizeof(int32_t)))) return st;
    abea::RecalArgs a{};
    a.rec = r->rec.as<abea_event_alignment_t>();
    a.n_rec = num_alignments;
    a.events = r->events.as<abea_event_t>();
    a.n_events = (int32_t)std::min<size_t>(et.n, 0x7fffffff);
    a.model = r->model.as<abea_model_t>();
    a.scal = r->scal.as<abea_scalings_t>();
    a.status = r->status.as<int32_t>();
    a.scale_var = scale_var ? 1 : 0;
    HIP_TRY(abea::launch_recalibrate_records(a, s));
    return ABEA_OK;
  };
  if (go()) return false;
  int32_t status = 0;
  abea_scalings_t sc;
  if (download(&status, r->status, 1, s) || download(&sc, r->scal, 1, s)) return false;
  if (hipStreamSynchronize(s) != hipSuccess) {
    fail(ABEA_ERR_HIP, "recalibrate_model: stream synchronize failed");
    return false;
  }
  if (status < 0) {
    fail(ABEA_ERR_INVALID, "recalibrate_model: an 'M' record's event_idx is outside the event table");
    return false;
  }
  if (status == 0) return false;
  *scallings = sc;
  return true;
}

abea_status_t abea_postalign_db(int32_t n_bam_rec, char **read, const int32_t *read_len,
                                const abea_event_table *et, const abea_model_t *model,
                                abea_aligned_pair_t *const *event_align_pairs,
                                const int32_t *n_event_align_pairs, abea_scalings_t *scalings,
                                abea_event_alignment_t **event_alignment,
                                int32_t *n_event_alignment, double *events_per_base,
                                abea_index_pair_t **base_to_event_map, int32_t *read_stat_flag) {
  if (n_bam_rec < 0 ||
      (n_bam_rec > 0 && (!read || !read_len || !et || !model || !event_align_pairs ||
                         !n_event_align_pairs || !scalings || !event_alignment ||
                         !n_event_alignment || !events_per_base || !base_to_event_map ||
                         !read_stat_flag)))
    return fail(ABEA_ERR_INVALID, "postalign_db: null argument");
  if (n_bam_rec == 0) return ABEA_OK;
  for (int32_t i = 0; i < n_bam_rec; ++i)
    if (n_event_align_pairs[i] > 0 && !event_align_pairs[i])
      return fail(ABEA_ERR_INVALID, "postalign_db: read %d has pairs but a null array", i);
  HostLease lease;
  abea_status_t st = lease.acquire();
  if (st) return st;
  HostRun *r = lease.r;
  HIP_TRY(hipSetDevice(r->ctx->device));
  HostPost h;
  host_post_layout(h, n_bam_rec, read, read_len, et, n_event_align_pairs);
  std::vector<int32_t> flag, n_rec;
  std::vector<double> epb;
  if ((st = run_postalign(r, h, read, read_len, et, model, event_align_pairs, n_event_align_pairs,
                          scalings, 1, flag, n_rec, epb)))
    return st;
  std::vector<abea_event_alignment_t> rec_h;
  std::vector<abea_index_pair_t> bmap_h;
  int64_t nr = 0, nb = 0;
  for (int32_t i = 0; i < n_bam_rec; ++i) {
    nr += std::max(n_event_align_pairs[i], 0);
    nb += std::max(h.meta[i].n_k, 0);
  }
  rec_h.resize((size_t)nr);
  bmap_h.resize((size_t)nb);
  hipStream_t s = r->ctx->stream;
  if ((st = download(rec_h.data(), r->rec, rec_h.size(), s))) return st;
  if ((st = download(bmap_h.data(), r->bmap, bmap_h.size(), s))) return st;
  HIP_TRY(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n_bam_rec; ++i) {
    read_stat_flag[i] = flag[i];
    event_alignment[i] = nullptr;
    base_to_event_map[i] = nullptr;
    n_event_alignment[i] = 0;
    events_per_base[i] = 0.0;
    if (n_event_align_pairs[i] <= 0) continue;
    const ReadMeta &m = h.meta[i];
    // f5c allocates n_event_align_pairs records and n_kmers map entries
    event_alignment[i] = (abea_event_alignment_t *)malloc(sizeof(abea_event_alignment_t) *
                                                          (size_t)n_event_align_pairs[i]);
    base_to_event_map[i] =
        (abea_index_pair_t *)malloc(sizeof(abea_index_pair_t) * (size_t)std::max(m.n_k, 1));
    if (!event_alignment[i] || !base_to_event_map[i]) {
      // the caller owns nothing from a failed call
      for (int32_t j = 0; j <= i; ++j) {
        free(event_alignment[j]);
        free(base_to_event_map[j]);
      }
      for (int32_t j = 0; j < n_bam_rec; ++j) {
        event_alignment[j] = nullptr;
        base_to_event_map[j] = nullptr;
        n_event_alignment[j] = 0;
        events_per_base[j] = 0.0;
      }
      return fail(ABEA_ERR_NOMEM, "postalign_db: malloc failed");
    }
    std::memcpy(event_alignment[i], rec_h.data() + m.pair_off,
                sizeof(abea_event_alignment_t) * (size_t)n_rec[i]);
    std::memcpy(base_to_event_map[i], bmap_h.data() + m.bmap_off,
                sizeof(abea_index_pair_t) * (size_t)m.n_k);
    n_event_alignment[i] = n_rec[i];
    events_per_base[i] = epb[i];
  }
  return ABEA_OK;
}

}  // extern "C"
