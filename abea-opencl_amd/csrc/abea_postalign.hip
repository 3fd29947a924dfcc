// Post-align kernels for gfx950 (CDNA4, wave64), one wavefront per read:
//
//   postalign    f5c postalign() then recalibrate_model() and the read
//                status (GPU/f5cmisc.h:51-54, flags GPU/f5c.h:49-51) from
//                align()'s AlignedPairs
//   recal_records recalibrate_model() from event alignment records
//   mom          estimate_scalings_using_mom() (GPU/f5cmisc.h:47)
//
// The reference only declares these functions; their bodies are f5c's ports
// of nanopolish.  What they compute:
//   postalign: each event on the path belongs to the k-mer of its first pair;
//     base_to_event_map[k] = {first, last} event of k-mer k or {-1, -1};
//     events_per_base = (last event - first event) / n_kmers; one
//     event_alignment_t per event, in k-mer (= event) order, hmm_state 'M'
//     when the k-mer rank differs from the previous record's, else 'E'.
//   recalibrate_model: weighted least squares of event mean on level mean,
//     weights 1/level_stdv^2, over the 'M' records, solved by Cramer's rule;
//     var = sqrt(mean squared weighted residual); at least 200 'M' records.
//   MOM: shift = mean(event means) - mean(levels); scale = mean((x -
//     shift)^2) / mean(level^2); var = log_var = 0.
//
// postalign in parallel.  On an alignment path each step goes to the next
// event (U), the next k-mer (L) or both (D).  Call pair p "new" when it is
// the first pair of its event (p = 0 or e_p != e_{p-1}).  Then, from the
// pair's neighbours alone:
//   e_p is k_p's first event  unless pair p-1 is a new pair on k-mer k_p
//   e_p is k_p's last event   unless pair p+1 is a new pair on k-mer k_p
//   k_p has no event          when p is not new (entered by L) and pair p+1
//                             is on another k-mer
// and the records are the new pairs in path order (record e_p - e_0).  The
// previous record's k-mer, for hmm_state, is the largest k-mer of the new
// pairs before p: a DPP max-scan per 64 pairs with a carry.  Every map entry
// is written once, so no ordering between lanes or chunks is needed.
//
// Sums.  The least-squares and moment sums are serial double sums in record
// (event, k-mer) order whose results are rounded to float.  Lanes form the
// terms in parallel and stage them in LDS; the sums add them back in order,
// so they round exactly as the serial CPU loop does.  Division and
// square root are the correctly rounded IEEE operations; log(var) comes from
// the device math library (<= 1 ulp in double before rounding to float).
#include <hip/hip_runtime.h>

#include "abea_internal.h"
#include "abea_wave.h"

namespace abea {
namespace {

constexpr int kMinMRecords = 200;     // minNumEventsToRescale of recalibrate_model
constexpr float kMinCalibrationVar = ABEA_MIN_CALIBRATION_VAR;
constexpr double kMaxEventsPerBase = ABEA_MAX_EVENTS_PER_BASE;
constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ float event_mean(const abea_event_t *ev, const float *evm, int64_t i) {
  return evm ? evm[i] : ev[i].mean;
}

// sums of the weighted least-squares normal equations
struct Normal {
  double a00 = 0, a01 = 0, a11 = 0, b0 = 0, b1 = 0;
};

// In-order sums.  Each chunk's terms go to the wave's LDS stage, compacted
// to the contributing lanes in lane order; the sums then read them back in
// that order (every lane reads the same address: a broadcast) and add them
// one at a time, as the serial CPU loop does.
constexpr int kStage = 6;  // doubles per staged record (48 B keeps 16-byte alignment)

struct alignas(16) Stage {
  double t[64 * kStage];
};

// One 'M' record's terms (x = event mean, level mean mu, level stdv sd).
struct Terms {
  double w, mw, mmw, xw, mxw;
};

__device__ __forceinline__ Terms normal_terms(float xf, float muf, float sdf) {
  const double x = xf, mu = muf, sd = sdf;
  Terms t;
  t.w = 1.0 / (sd * sd);
  t.mw = mu * t.w;
  t.mmw = mu * mu * t.w;
  t.xw = x * t.w;
  t.mxw = mu * x * t.w;
  return t;
}

// Stage the terms of the lanes with m set (record order = lane order) and
// add them to the normal-equation sums; returns how many there were.
__device__ __forceinline__ int add_terms(Stage &st, bool m, const Terms &t, Normal &n) {
  const uint64_t mask = __builtin_amdgcn_ballot_w64(m);
  if (m) {
    double *q = st.t + kStage * lanes_below(mask);
    *reinterpret_cast<double2 *>(q) = make_double2(t.w, t.mw);
    *reinterpret_cast<double2 *>(q + 2) = make_double2(t.mmw, t.xw);
    q[4] = t.mxw;
  }
  wave_lds_sync();
  const int cnt = __builtin_popcountll(mask);
  for (int j = 0; j < cnt; ++j) {
    const double *q = st.t + kStage * j;
    const double2 p0 = *reinterpret_cast<const double2 *>(q);
    const double2 p1 = *reinterpret_cast<const double2 *>(q + 2);
    n.a00 += p0.x;
    n.a01 += p0.y;
    n.a11 += p1.x;
    n.b0 += p1.y;
    n.b1 += q[4];
  }
  wave_lds_sync();
  return cnt;
}

// acc += v[0] + v[S] + v[2S] + ... over n entries (n <= 64), one add at a
// time in order, the LDS loads of 16 entries issued ahead of their adds.
// Entries n .. (n rounded up to 16) - 1 must hold +0.0, which leaves acc
// unchanged: acc starts at +0 and is never -0 (a sum that cancels exactly
// rounds to +0).
template <int S>
__device__ __forceinline__ void serial_sum_lds(const double *v, int n, double &acc) {
  for (int j0 = 0; j0 < n; j0 += 16) {
    double t[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) t[j] = v[S * (j0 + j)];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc += t[j];
  }
}

// One running sum over the lanes with m set, in lane order.
__device__ __forceinline__ void add_in_order(Stage &st, bool m, double t, double &acc) {
  const uint64_t mask = __builtin_amdgcn_ballot_w64(m);
  if (m) st.t[lanes_below(mask)] = t;
  wave_lds_sync();
  const int cnt = __builtin_popcountll(mask);
  for (int j = 0; j < cnt; ++j) acc += st.t[j];
  wave_lds_sync();
}

__device__ __forceinline__ double var_term(float xf, float muf, float sdf, double shift,
                                           double scale) {
  const double x = xf, mu = muf, sd = sdf;
  const double y = x - shift - scale * mu;
  return y * y / (sd * sd);
}

// Cramer's rule on the 2x2 normal equations: (shift, scale).
__device__ __forceinline__ void solve(const Normal &n, double &shift, double &scale) {
  const double a10 = n.a01;
  const double det = n.a00 * n.a11 - n.a01 * a10;
  shift = -(n.a01 * n.b1 - n.a11 * n.b0) / det;
  scale = (n.a00 * n.b1 - a10 * n.b0) / det;
}

__device__ __forceinline__ abea_scalings_t finish(double shift, double scale, double var) {
  abea_scalings_t o;
  o.scale = (float)scale;
  o.shift = (float)shift;
  o.var = (float)var;
  o.log_var = (float)log(var);
  return o;
}

// ---------------------------------------------------------------------------
// postalign + recalibrate_model from pairs: a walk kernel over segments of
// each read's pairs, then a finish kernel per read.

constexpr int kSegPairs = 2048;  // pairs per walk work item (32 chunks)

struct PairView {
  const abea_aligned_pair_t *P;
  int32_t np, nk, n_e;
  const char *S;
};

// Per-lane decisions for pair p of a chunk.
struct PairStep {
  int k, e;
  bool valid;  // p inside the segment
  bool bad;    // not a step of an alignment path, or out of range
  bool is_new, first, last, none;
  bool m;      // the pair's record is an 'M' record
  uint32_t lo, hi;  // the k-mer's bases 0-3 and 4-5
  uint32_t rank;
};

// What one chunk hands to the next: its last pair, whether that pair was
// new, and the rank of the last record so far (-1: none yet).
struct Carry {
  int k = -1, e = -1, is_new = 0, rank = -1;
};

__device__ __forceinline__ int2 load_pair(const abea_aligned_pair_t *P, int i) {
  const int2 *q = reinterpret_cast<const int2 *>(P);
  return q[i];
}

// A chunk's raw loads: its pairs and each pair's k-mer bases.  Loads are
// unconditional, from clamped (in-bounds) indices, so that they can be
// issued ahead: a load under a branch would be waited for at the branch's
// join.  Validity is applied where they are used.
struct ChunkLoad {
  int2 c;
  uint32_t lo, hi4;  // bases k..k+3 and k+2..k+5 of each pair's k-mer k
};

// A chunk's gathered recalibration inputs: each lane's event mean and its
// k-mer's model row.
struct ChunkGather {
  float x = 0.f, mu = 0.f, sd = 0.f;
};

__device__ __forceinline__ void load_pairs(const PairView &v, int base, int lane, ChunkLoad &L) {
  L.c = load_pair(v.P, min(base + lane, v.np - 1));
}

// Two 4-byte loads inside the read (k <= len - 6), with no sub-word load
// whose widening would be done, and waited for, as soon as it is issued.
__device__ __forceinline__ void load_kmers(const PairView &v, ChunkLoad &L) {
  const char *q = v.S + min(max(L.c.x, 0), v.nk - 1);
  __builtin_memcpy(&L.lo, q, 4);
  __builtin_memcpy(&L.hi4, q + 2, 4);
}

// All lanes must call.  Pairs p < end belong to this work item; the path's
// pair p+1 exists when p + 1 < np.  Neighbouring pairs come from the
// neighbouring lanes (DPP lane shifts); lane 0 takes pair p-1 from the carry,
// lane 63 takes pair p+1 from nx (the next chunk's first pair, clamped to the
// path's last as the loads are).  A pair is bad when its step from pair p-1
// is not D, U or L, its k-mer or event is out of range, or its record index
// e - e0 is outside [0, np).
__device__ __forceinline__ PairStep pair_step(const PairView &v, int base, int lane, int end,
                                              int e0, const ChunkLoad &L, int2 nx,
                                              bool need_events, Carry &cy) {
  PairStep s{};
  const int p = base + lane;
  s.valid = p < end;
  // neighbours from the loaded pairs themselves: pair p+1 may lie past this
  // work item's end (it belongs to the next one), but it is the path's
  const int2 c = s.valid ? L.c : make_int2(-1, -1);
  const int nk_ = __builtin_amdgcn_update_dpp(nx.x, L.c.x, 0x130, 0xf, 0xf, false);  // wave_shl:1
  const int ne_ = __builtin_amdgcn_update_dpp(nx.y, L.c.y, 0x130, 0xf, 0xf, false);
  const int ak = __builtin_amdgcn_update_dpp(cy.k, L.c.x, 0x138, 0xf, 0xf, false);  // wave_shr:1
  const int ae = __builtin_amdgcn_update_dpp(cy.e, L.c.y, 0x138, 0xf, 0xf, false);
  s.k = c.x;
  s.e = c.y;
  bool ok = c.x >= 0 && c.x < v.nk && c.y >= 0 && (!need_events || c.y < v.n_e) &&
            c.y - e0 >= 0 && c.y - e0 < v.np;
  if (p >= 1) {
    const int dk = c.x - ak, de = c.y - ae;
    ok = ok && (dk == 0 || dk == 1) && (de == 0 || de == 1) && (dk | de) != 0;
  }
  s.bad = s.valid && !ok;
  const bool live = s.valid && ok;
  s.is_new = live && (p == 0 || c.y != ae);
  const bool prev_new =
      __builtin_amdgcn_update_dpp(cy.is_new, (int)s.is_new, 0x138, 0xf, 0xf, false) != 0;
  s.first = s.is_new && !(p >= 1 && ak == c.x && prev_new);
  s.last = s.is_new && !(p + 1 < v.np && ne_ != c.y && nk_ == c.x);
  s.none = live && !s.is_new && (p + 1 == v.np || nk_ != c.x);
  s.lo = live ? L.lo : 0;
  s.hi = live ? L.hi4 >> 16 : 0;
  s.rank = live ? rank_lohi(s.lo, s.hi) : 0;

  // rank of the previous record: that of the last new pair before this one.
  // Ranks are below 4096, so the prefix maximum of lane * 4096 + rank over
  // the new pairs finds the last new lane at or before each lane and carries
  // its rank in the low 12 bits.
  const int incl = wave_scan_max(s.is_new ? (lane << 12) | (int)s.rank : -1);
  const int excl = __builtin_amdgcn_update_dpp(-1, incl, 0x138, 0xf, 0xf, false);  // wave_shr:1
  const int prank = excl >= 0 ? (excl & 0xfff) : cy.rank;
  s.m = s.first && (prank < 0 || (uint32_t)prank != s.rank);

  const int last_new = __builtin_amdgcn_readlane(incl, 63);
  if (last_new >= 0) cy.rank = last_new & 0xfff;
  cy.k = __builtin_amdgcn_readlane(L.c.x, 63);
  cy.e = __builtin_amdgcn_readlane(L.c.y, 63);
  cy.is_new = __builtin_amdgcn_readlane((int)s.is_new, 63);
  return s;
}

// The 32 bytes of a pair's event_alignment_t record (x86-64 layout).
__device__ __forceinline__ void record_words(const PairStep &s, uint4 &w0, uint4 &w1) {
  w0.x = s.lo;                          // ref_kmer[0..3]
  w0.y = s.hi;                          // ref_kmer[4..5], NUL, pad
  w0.z = (uint32_t)s.k;                 // ref_position
  w0.w = 0xffffffffu;                   // read_idx = -1
  w1.x = (uint32_t)s.e;                 // event_idx
  w1.y = s.lo << 8;                     // rc = 0, model_kmer[0..2]
  w1.z = (s.lo >> 24) | (s.hi << 8);    // model_kmer[3..5], NUL
  w1.w = s.m ? 'M' : 'E';               // hmm_state, pad
}

// Stores that every lane of a chunk issues: raw buffer stores (vector
// memory) whose lanes with nothing to write pass a byte offset outside the
// buffer's range, which the hardware drops.  Each chunk then issues the same
// number of memory operations whichever lanes write, so the compiler can
// wait for a load by count without also waiting for younger stores.
constexpr uint32_t kSkip = 0x80000000u;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// A raw buffer of `bytes` bytes at p (descriptor word 3: 32-bit data format,
// the gfx9-family value for untyped buffer access).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t store_buffer(void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void store_b32(uint32_t v, __amdgpu_buffer_rsrc_t b, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(v, b, (int)off, 0, 0);
}
__device__ __forceinline__ void store_b64(u32x2 v, __amdgpu_buffer_rsrc_t b, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b64(v, b, (int)off, 0, 0);
}
__device__ __forceinline__ void store_b128(u32x4 v, __amdgpu_buffer_rsrc_t b, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b128(v, b, (int)off, 0, 0);
}

// A chunk's records are consecutive (its new pairs have consecutive events):
// assemble them in the wave's LDS stage L (128 words of 16 bytes) and store
// them as one contiguous block, 16 bytes per lane per store, into the read's
// records (buffer rb, record e - e0 of event e).  A chunk without new pairs
// issues the same two stores, all lanes skipped.  The records are output
// only: the stores are marked non-temporal (streamed past the caches).
constexpr int kStoreNT = 2;  // cache-policy bits of a buffer store: nt
__device__ __forceinline__ void write_records(uint4 *L, __amdgpu_buffer_rsrc_t rb, int e0,
                                              const PairStep &s, int lane) {
  const uint64_t nm = __builtin_amdgcn_ballot_w64(s.is_new);
  if (s.is_new) {
    uint4 w0, w1;
    record_words(s, w0, w1);
    const int j = lanes_below(nm);
    L[2 * j] = w0;
    L[2 * j + 1] = w1;
  }
  wave_lds_sync();
  const int e_first = nm ? __builtin_amdgcn_readlane(s.e, __builtin_ctzll(nm)) : e0;
  const uint32_t o = 32u * (uint32_t)(e_first - e0);
  const int words = 2 * __builtin_popcountll(nm);
  const uint4 q0 = L[lane], q1 = L[64 + lane];
  __builtin_amdgcn_raw_buffer_store_b128(u32x4{q0.x, q0.y, q0.z, q0.w}, rb,
                                         (int)(lane < words ? o + 16u * lane : kSkip), 0,
                                         kStoreNT);
  __builtin_amdgcn_raw_buffer_store_b128(u32x4{q1.x, q1.y, q1.z, q1.w}, rb,
                                         (int)(64 + lane < words ? o + 16u * (64 + lane) : kSkip),
                                         0, kStoreNT);
  wave_lds_sync();
}

// The rank of the record of the last new pair at or before pair q (q >= 0):
// step back over pairs that repeat their predecessor's event.
__device__ __forceinline__ int last_record_rank(const PairView &v, int q) {
  int2 c = load_pair(v.P, q);
  while (q > 0) {
    const int2 b = load_pair(v.P, q - 1);
    if (b.y != c.y) break;
    c = b;
    --q;
  }
  uint32_t lo, hi;
  kmer_bytes(v.S, min(max(c.x, 0), v.nk - 1), lo, hi);
  return (int)rank_lohi(lo, hi);
}

// Walk kernel: one wave per (read, segment of kSegPairs pairs), kWalkWaves
// waves per workgroup.  Writes the map entries (MAP), the records (REC) and,
// per k-mer on the path (RECAL), the (event mean, level mean, level stdv,
// 'M') of the k-mer's first record -- each k-mer of the path has exactly one
// pair that writes it: the first pair of its events or, for a k-mer without
// events, the pair that passes it.  A bad path sets the read's status word.
// With RECAL the workgroup first copies the model's level means and stdvs
// (32 KB) into LDS, where the pairs' k-mer ranks gather them: gathered from
// global memory, 64 scattered rows per chunk cost more than the rest of the
// chunk's loads together.
// Eight waves share a table: 48 KB of LDS per workgroup with the record
// stages, three workgroups (24 waves) per CU at up to 80 VGPRs per wave.
constexpr int kWalkWaves = 8;
template <bool REC, bool MAP, bool RECAL>
__global__ __launch_bounds__(64 * kWalkWaves, 6) void abea_postalign_walk_kernel(PostArgs a) {
  __shared__ uint4 rec_stage[REC ? kWalkWaves : 1][128];
  __shared__ float2 levels[RECAL ? ABEA_NUM_KMER : 1];
  if constexpr (RECAL) {
    for (int i = threadIdx.x; i < ABEA_NUM_KMER; i += 64 * kWalkWaves)
      levels[i] = make_float2(a.model[i].level_mean, a.model[i].level_stdv);
    __syncthreads();
  }
  uint4 *st = rec_stage[REC ? threadIdx.x >> 6 : 0];
  const int lane = threadIdx.x & 63;
  const int w = uniform((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (w >= a.n_items) return;
  const int2 item = a.items[w];
  const int r = uniform(item.x);
  const int begin = uniform(item.y) * kSegPairs;
  const int np = uniform(a.n_pairs[r]);
  if (begin >= np) return;
  const ReadMeta m = a.meta[r];
  PairView v;
  v.P = a.pairs + m.pair_off;
  v.np = np;
  v.nk = uniform(m.n_k);
  v.n_e = uniform(m.n_e);
  v.S = a.seq + a.seq_ptr[r];
  // pairs on a read without k-mers (or, to recalibrate, without events)
  // cannot be a path
  if (v.nk <= 0 || (RECAL && v.n_e <= 0)) {
    if (lane == 0) a.pstat[r] = 1;
    return;
  }
  const int end = min(begin + kSegPairs, np);
  const int64_t evb = m.ev_off;
  const __amdgpu_buffer_rsrc_t rb =
      store_buffer(REC ? a.rec + m.pair_off : nullptr, REC ? 32u * (uint32_t)np : 0u);
  const __amdgpu_buffer_rsrc_t mb =
      store_buffer(MAP ? a.bmap + m.bmap_off : nullptr, MAP ? 8u * (uint32_t)v.nk : 0u);
  const __amdgpu_buffer_rsrc_t tb =
      store_buffer(RECAL ? a.terms + m.bmap_off : nullptr, RECAL ? 16u * (uint32_t)v.nk : 0u);
  const int2 p0 = load_pair(v.P, 0);
  const int e0 = uniform(p0.y);

  if (MAP && begin == 0) {  // k-mers before the path's first and after its last
    int2 *B2 = reinterpret_cast<int2 *>(a.bmap + m.bmap_off);
    const int k0 = uniform(p0.x), kN = uniform(load_pair(v.P, np - 1).x);
    const int2 none = make_int2(-1, -1);
    for (int k = lane; k < min(k0, v.nk); k += 64) B2[k] = none;
    for (int k = max(kN + 1, 0) + lane; k < v.nk; k += 64) B2[k] = none;
  }

  Carry cy;
  if (begin > 0) {
    const int2 q1 = load_pair(v.P, begin - 1);
    cy.k = uniform(q1.x);
    cy.e = uniform(q1.y);
    cy.is_new = begin - 1 == 0 || uniform(load_pair(v.P, begin - 2).y) != cy.e;
    cy.rank = uniform(last_record_rank(v, begin - 1));
  }

  // Software pipeline: each load level waits for the one before it (pairs,
  // then their k-mer bases, then the gathers of model rows and event means),
  // so each runs one chunk ahead of the next.  The step for chunk c issues
  // the gathers of chunk c+1, the pair loads of chunk c+3 and the k-mer
  // loads of chunk c+2, then works on chunk c and issues its stores.  Every
  // load it waits for was issued before the previous chunk's work, and with
  // the same memory operations in every step (the stores above) and a loop
  // with one exit, the waits count exactly and never wait for the previous
  // chunk's stores.  The loop runs whole groups of four steps, the buffers'
  // roles rotated so that no buffer is copied (a copy of a register still
  // being loaded waits for the load); steps past the segment's end have no
  // valid pairs and store nothing.  After a bad pair the read's walk stops
  // storing (dead) and its status word is set at the end.
  auto gather = [&](const ChunkLoad &L, ChunkGather &g) {
    if constexpr (RECAL) {
      const float2 md = levels[rank_lohi(L.lo, L.hi4 >> 16)];
      g.mu = md.x;
      g.sd = md.y;
      g.x = event_mean(a.events, a.ev_mean, evb + min(max(L.c.y, 0), v.n_e - 1));
    }
  };
  // The stores of one chunk; lanes whose flags are clear store nothing.
  auto stores = [&](const PairStep &s, const ChunkGather &g) {
    if constexpr (REC) write_records(st, rb, e0, s, lane);
    if constexpr (MAP) {
      // {first, last} event, or {-1, -1} for a k-mer without events, as one
      // entry; a first or last event alone as one word
      const bool whole = (s.first && s.last) || s.none;
      const uint32_t ev = (uint32_t)s.e;
      store_b64(s.none ? u32x2{~0u, ~0u} : u32x2{ev, ev}, mb, whole ? 8u * s.k : kSkip);
      store_b32(ev, mb, s.first != s.last ? 8u * s.k + (s.last ? 4u : 0u) : kSkip);
    }
    if constexpr (RECAL) {
      const u32x4 t = s.m ? u32x4{__builtin_bit_cast(uint32_t, g.x),
                                  __builtin_bit_cast(uint32_t, g.mu),
                                  __builtin_bit_cast(uint32_t, g.sd), 0x3f800000u}
                          : u32x4{0u, 0u, 0u, 0u};
      store_b128(t, tb, (s.first || s.none) ? 16u * s.k : kSkip);
    }
  };
  bool dead = false;
  // Chunk `base` (loads in A, gathers in GA; B, C, D the next three chunks).
  auto step = [&](int base, const ChunkLoad &A, ChunkLoad &B, ChunkLoad &C, ChunkLoad &D,
                  const ChunkGather &GA, ChunkGather &GB) {
    // keep each step's work inside the step: the scheduler would otherwise
    // start the next step's arithmetic on loads issued moments before
    __builtin_amdgcn_sched_barrier(0);
    gather(B, GB);
    load_pairs(v, base + 192, lane, D);
    load_kmers(v, C);
    // the next chunk's first pair (lane 0 loads min(base + 64, np - 1))
    const int2 nx =
        make_int2(__builtin_amdgcn_readlane(B.c.x, 0), __builtin_amdgcn_readlane(B.c.y, 0));
    PairStep s = pair_step(v, base, lane, end, e0, A, nx, RECAL, cy);
    dead = dead || __builtin_amdgcn_ballot_w64(s.bad) != 0;
    s.is_new = s.is_new && !dead;
    s.first = s.first && !dead;
    s.last = s.last && !dead;
    s.none = s.none && !dead;
    s.m = s.m && !dead;
    stores(s, GA);
  };
  ChunkLoad c0, c1, c2, c3;
  ChunkGather g0, g1;
  load_pairs(v, begin, lane, c0);
  load_pairs(v, begin + 64, lane, c1);
  load_kmers(v, c0);
  gather(c0, g0);
  load_pairs(v, begin + 128, lane, c2);
  load_kmers(v, c1);
  stores(PairStep{}, g0);  // the memory operations a step leaves behind
  for (int base = begin; base < end; base += 256) {
    step(base, c0, c1, c2, c3, g0, g1);
    step(base + 64, c1, c2, c3, c0, g1, g0);
    step(base + 128, c2, c3, c0, c1, g0, g1);
    step(base + 192, c3, c0, c1, c2, g1, g0);
  }
  if (dead && lane == 0) a.pstat[r] = 1;
}

// The five normal-equation sums over the 'M' entries of T[k0..kN] (the
// k-mers' first records, in record order).  Each chunk's terms are staged in
// LDS, compacted to the 'M' entries, one record per 48 bytes; lane c < 5 then
// runs sum c over them in order (one LDS read and one add per record, the
// other lanes masked off), and lanes 0..4 hold the five sums at the end.
// n_m receives the count.
//
// Both passes keep the entries of the next three chunks loading: three
// buffers, each reloaded with the chunk three ahead as soon as its own
// chunk's terms are formed, in whole groups of three chunks (chunks past kN
// have no entries), so that no buffer is copied -- a copy of a register
// still being loaded waits for the load.
__device__ __forceinline__ Normal normal_sums(Stage &st, const float4 *T, int k0, int kN,
                                              int lane, int &n_m) {
  double acc = 0.0;
  n_m = 0;
  auto step = [&](int base, float4 &q) {
    const bool mm = base + lane <= kN && q.w != 0.f;
    const uint64_t mask = __builtin_amdgcn_ballot_w64(mm);
    const int cnt = __builtin_popcountll(mask);
    const Terms t = normal_terms(q.x, q.y, q.z);
    // q is dead from here: the reload may take its registers (issued any
    // earlier, it would need others, and the loop would copy them back)
    __builtin_amdgcn_sched_barrier(0);
    q = T[min(base + 192 + lane, kN)];
    if (lane >= cnt) {  // padding for the 16-entry reads (positions cnt..63)
      double *z = st.t + kStage * lane;
      *reinterpret_cast<double2 *>(z) = make_double2(0.0, 0.0);
      *reinterpret_cast<double2 *>(z + 2) = make_double2(0.0, 0.0);
      z[4] = 0.0;
    }
    if (mm) {  // the entries, compacted to positions 0..cnt-1
      double *w = st.t + kStage * lanes_below(mask);
      *reinterpret_cast<double2 *>(w) = make_double2(t.w, t.mw);
      *reinterpret_cast<double2 *>(w + 2) = make_double2(t.mmw, t.xw);
      w[4] = t.mxw;
    }
    wave_lds_sync();
    if (lane < 5) serial_sum_lds<kStage>(st.t + lane, cnt, acc);
    wave_lds_sync();
    n_m += cnt;
  };
  float4 b0 = T[min(k0 + lane, kN)], b1 = T[min(k0 + 64 + lane, kN)],
         b2 = T[min(k0 + 128 + lane, kN)];
  for (int base = k0; base <= kN; base += 192) {
    step(base, b0);
    step(base + 64, b1);
    step(base + 128, b2);
  }
  Normal nm;
  nm.a00 = readlane_f64(acc, 0);
  nm.a01 = readlane_f64(acc, 1);
  nm.a11 = readlane_f64(acc, 2);
  nm.b0 = readlane_f64(acc, 3);
  nm.b1 = readlane_f64(acc, 4);
  return nm;
}

// Sum of the squared weighted residuals over the same entries, in order
// (lane 0 runs it), broadcast.
__device__ __forceinline__ double var_sum(Stage &st, const float4 *T, int k0, int kN, int lane,
                                          double shift, double scale) {
  double acc = 0.0;
  auto step = [&](int base, float4 &q) {
    const bool mm = base + lane <= kN && q.w != 0.f;
    const uint64_t mask = __builtin_amdgcn_ballot_w64(mm);
    const int cnt = __builtin_popcountll(mask);
    const double t = var_term(q.x, q.y, q.z, shift, scale);
    __builtin_amdgcn_sched_barrier(0);  // q is dead from here (see normal_sums)
    q = T[min(base + 192 + lane, kN)];
    if (lane >= cnt) st.t[lane] = 0.0;  // padding for the 16-entry reads
    if (mm) st.t[lanes_below(mask)] = t;  // the entries, compacted
    wave_lds_sync();
    if (lane == 0) serial_sum_lds<1>(st.t, cnt, acc);
    wave_lds_sync();
  };
  float4 b0 = T[min(k0 + lane, kN)], b1 = T[min(k0 + 64 + lane, kN)],
         b2 = T[min(k0 + 128 + lane, kN)];
  for (int base = k0; base <= kN; base += 192) {
    step(base, b0);
    step(base + 64, b1);
    step(base + 128, b2);
  }
  return readlane_f64(acc, 0);
}

// Finish kernel: one wave per read.  Counts and status, then (with
// recalibration) the in-order sums, the solve and the variance.
__global__ __launch_bounds__(256) void abea_postalign_finish_kernel(PostArgs a) {
  __shared__ Stage stages[kWavesPerBlock];
  Stage &st = stages[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int w = uniform((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (w >= a.n_reads) return;
  const int r = uniform(a.order ? a.order[w] : w);
  const abea_scalings_t sc = a.scal_in[r];
  const int np = uniform(a.n_pairs[r]);
  if (np <= 0) {
    if (lane == 0) {
      a.n_rec[r] = 0;
      if (a.epb) a.epb[r] = 0.0;
      a.scal_out[r] = sc;
      a.flag[r] = ABEA_FAILED_ALIGNMENT;
    }
    return;
  }
  // the walk's verdict (a vector load: the walk kernel stored it)
  if (uniform(*(volatile int32_t *)&a.pstat[r]) != 0) {
    if (lane == 0) {
      a.pstat[r] = 0;
      a.n_rec[r] = 0;
      a.scal_out[r] = sc;
      a.flag[r] = kBadPath;
    }
    return;
  }
  const ReadMeta m = a.meta[r];
  const abea_aligned_pair_t *P = a.pairs + m.pair_off;
  const int2 p0 = load_pair(P, 0), pN = load_pair(P, np - 1);
  const int e0 = uniform(p0.y), eN = uniform(pN.y);
  const int k0 = uniform(p0.x), kN = uniform(pN.x);
  const int nk = uniform(m.n_k);

  abea_scalings_t out = sc;
  int32_t flag = 0;
  if (a.recalibrate) {
    const float4 *T = a.terms + m.bmap_off;
    int n_m = 0;
    const Normal nm = normal_sums(st, T, k0, kN, lane, n_m);
    if (n_m >= kMinMRecords) {
      double shift, scale;
      solve(nm, shift, scale);
      double var = 1.0;
      if (a.scale_var) {
        var = var_sum(st, T, k0, kN, lane, shift, scale);
        var /= n_m;
        var = sqrt(var);
      }
      out = finish(shift, scale, var);
      if (out.var > kMinCalibrationVar) flag = ABEA_FAILED_CALIBRATION;
    } else {
      flag = ABEA_FAILED_CALIBRATION;
    }
  }
  const double epb = (double)(eN - e0) / nk;
  // f5c's stay filter after a passed calibration: too many events per base
  // fails the quality check (FAILED_QUALITY_CHK, GPU/f5c.h:52)
  if (a.recalibrate && flag == 0 && epb > kMaxEventsPerBase) flag = ABEA_FAILED_QUALITY_CHK;
  if (lane == 0) {
    a.n_rec[r] = eN - e0 + 1;
    if (a.epb) a.epb[r] = epb;
    a.scal_out[r] = out;
    a.flag[r] = flag;
  }
}

// ---------------------------------------------------------------------------
// recalibrate_model from event alignment records (a caller's own records).

struct RecView {
  bool valid, bad, m;
  float x;
  uint32_t rank;
};

__device__ __forceinline__ RecView rec_view(const RecalArgs &a, int i) {
  RecView v{};
  v.valid = i < a.n_rec;
  if (!v.valid) return v;
  const uint4 *q = reinterpret_cast<const uint4 *>(a.rec + i);
  const uint4 w0 = q[0], w1 = q[1];
  const int e = (int)w1.x;
  v.m = (w1.w & 0xff) == 'M';
  if (v.m) {
    if (e < 0 || e >= a.n_events) {
      v.bad = true;
      v.m = false;
      return v;
    }
    v.rank = rank_lohi(w0.x, w0.y & 0xffffu);
    v.x = a.events[e].mean;
  }
  return v;
}

__global__ __launch_bounds__(64) void abea_recal_records_kernel(RecalArgs a) {
  __shared__ Stage st;
  const int lane = threadIdx.x;
  int n_m = 0;
  bool bad = false;
  for (int base = 0; base < a.n_rec; base += 64) {
    const RecView v = rec_view(a, base + lane);
    if (__builtin_amdgcn_ballot_w64(v.bad)) {
      bad = true;
      break;
    }
    n_m += __builtin_popcountll(__builtin_amdgcn_ballot_w64(v.m));
  }
  if (bad) {
    if (lane == 0) *a.status = -1;
    return;
  }
  if (n_m < kMinMRecords) {
    if (lane == 0) *a.status = 0;
    return;
  }
  Normal nm;
  for (int base = 0; base < a.n_rec; base += 64) {
    const RecView v = rec_view(a, base + lane);
    Terms t{};
    if (v.m) {
      const abea_model_t md = a.model[v.rank];
      t = normal_terms(v.x, md.level_mean, md.level_stdv);
    }
    add_terms(st, v.m, t, nm);
  }
  double shift, scale;
  solve(nm, shift, scale);
  double var = 1.0;
  if (a.scale_var) {
    var = 0.0;
    for (int base = 0; base < a.n_rec; base += 64) {
      const RecView v = rec_view(a, base + lane);
      double t = 0.0;
      if (v.m) {
        const abea_model_t md = a.model[v.rank];
        t = var_term(v.x, md.level_mean, md.level_stdv, shift, scale);
      }
      add_in_order(st, v.m, t, var);
    }
    var /= n_m;
    var = sqrt(var);
  }
  if (lane == 0) {
    *a.scal = finish(shift, scale, var);
    *a.status = 1;
  }
}

// ---------------------------------------------------------------------------
// estimate_scalings_using_mom.

// Sum of float terms in parallel (each lane adds its own, then a tree), exact
// when provable: if every term is finite and a multiple of 2^g, and the sum of
// magnitudes stays below 2^(52+g) (one bit of margin for the rounding of that
// bound), every partial sum in any order is a double exactly, so the result
// equals the serial sum.
// The ulp exponent of a finite float grows with the bits of its magnitude, so
// a lane keeps the smallest and largest magnitude bits of its terms (zero
// counted as the largest unsigned value in the minimum) and converts them
// once: the smallest ulp exponent of the non-zero terms, and whether an inf
// or NaN (magnitude bits >= 0x7f800000) occurred.
struct ExactSum {
  double s = 0.0, bound = 0.0;  // this lane's partial sums
  uint32_t lo = 0xffffffffu, hi = 0u;
};

__device__ __forceinline__ void exact_add(ExactSum &es, float x) {
  const uint32_t u = __builtin_bit_cast(uint32_t, x) & 0x7fffffffu;
  es.lo = min(es.lo, u - 1u);  // zero wraps to the largest value
  es.hi = max(es.hi, u);
  es.s += (double)x;
  es.bound += fabs((double)x);
}

// The total (every lane's partial sums, tree-added); true when it provably
// equals the serial sum.
__device__ __forceinline__ bool exact_total(const ExactSum &es, double &total) {
  total = wave_sum_exact_f64(es.s);
  const double bound = wave_sum_exact_f64(es.bound);
  int g = 1 << 20;  // no non-zero term
  if (es.lo != 0xffffffffu) {
    const int eb = (int)((es.lo + 1u) >> 23);
    g = eb ? eb - 150 : -149;  // ulp exponent
  }
  const int gmin = -__builtin_amdgcn_readlane(wave_scan_max(-g), 63);
  const bool finite = __builtin_amdgcn_ballot_w64(es.hi >= 0x7f800000u) == 0;
  return finite && bound <= ldexp(1.0, min(52 + gmin, 1000));
}

// estimate_scalings_using_mom, one wave per read.  Pass 1 sums the event
// means and the k-mer levels (exact in parallel when provable, else again
// serially); pass 2 runs the two serial sums of squares side by side, the
// levels' in lane 0 and the events' in lane 1.  The serial chains of the
// longest read bound the launch, so the waves take the reads longest first
// (order) when the caller gives that order, and each pass keeps the loads of
// the next kMomGroup chunks of 64 terms in flight while it works on the
// current ones (one chunk ahead left a long read's wave waiting on memory
// for most of each chunk); the level means sit in LDS.
constexpr int kMomWaves = 8;
constexpr int kMomGroup = 4;
constexpr int kMomSpan = 64 * kMomGroup;  // terms per group
// the two chains' stages start 4 LDS banks apart (not 2 KB, the same bank),
// so the chain lanes' loads of one step do not conflict
constexpr int kMomSecond = kMomSpan + 2;

struct MomRead {
  const abea_event_t *events;
  const float *ev_mean;
  int64_t evb;
  const char *S;
  int n_e, nk, len;
};

struct MomLoad {
  float x;
  uint32_t lo, hi4;  // bases i..i+3 and i+2..i+5 of k-mer i
};

// Unconditional loads from clamped indices (see load_pairs), the same three
// in every group, so that waiting for one group does not wait for the next
// (the load counter is waited for in order, by count).  A read with no
// events or no k-mers reads a stand-in (the model table) for that kind; the
// values are not used.  The k-mer's bases come as two 4-byte loads inside
// the read (see load_kmers).
__device__ __forceinline__ void mom_load(const MomRead &d, int i, MomLoad &L) {
  L.x = event_mean(d.events, d.ev_mean, d.evb + max(min(i, d.n_e - 1), 0));
  const char *q = d.S + max(min(i, d.nk - 1), 0);
  __builtin_memcpy(&L.lo, q, 4);
  __builtin_memcpy(&L.hi4, q + 2, 4);
}

__device__ __forceinline__ void mom_load_group(const MomRead &d, int base, int lane,
                                               MomLoad (&L)[kMomGroup]) {
#pragma unroll
  for (int g = 0; g < kMomGroup; ++g) mom_load(d, base + 64 * g + lane, L[g]);
}

// acc += v[0] + v[1] + ... + v[kMomSpan - 1], one add at a time in index
// order, the LDS loads of 16 terms issued ahead of their adds.  The entries
// past a read's last term are +0.0, which leaves acc unchanged: acc starts
// at +0 and is never -0 (a sum that cancels exactly rounds to +0).
__device__ __forceinline__ void serial_sum_span(const double *v, double &acc) {
#pragma unroll 4
  for (int j0 = 0; j0 < kMomSpan; j0 += 16) {
    double t[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) t[j] = v[j0 + j];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc += t[j];
  }
}

__global__ __launch_bounds__(64 * kMomWaves) void abea_mom_kernel(MomArgs a) {
  __shared__ double stages[kMomWaves][kMomSecond + kMomSpan];
  __shared__ float levels[ABEA_NUM_KMER];
  for (int i = threadIdx.x; i < ABEA_NUM_KMER; i += blockDim.x) levels[i] = a.model[i].level_mean;
  __syncthreads();
  double *st = stages[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int q = uniform((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (q >= a.n_reads) return;
  const int r = a.order ? uniform(a.order[q]) : q;
  const ReadMeta m = a.meta[r];
  MomRead d;
  d.events = a.events;
  d.ev_mean = a.ev_mean;
  d.evb = m.ev_off;
  d.S = a.seq + a.seq_ptr[r];
  d.n_e = uniform(m.n_e);
  d.nk = uniform(m.n_k);
  d.len = d.nk + ABEA_KMER_SIZE - 1;
  if (d.n_e <= 0) {  // the loads' stand-ins (see mom_load)
    d.events = nullptr;
    d.ev_mean = &a.model[0].level_mean;
    d.evb = 0;
  }
  if (d.nk <= 0) d.S = reinterpret_cast<const char *>(a.model);
  const int n_e = d.n_e, nk = d.nk;
  const int n_max = max(n_e, nk);

  // Each pass runs over the read in whole pairs of groups (positions past
  // the read's last term are masked): two load buffers trade roles, one
  // group's loads in flight while the other group is used, and a buffer is
  // reloaded only after its last use (the scheduling barrier), so that no
  // buffer is copied -- a copy of a register still being loaded waits for
  // the load.
  MomLoad ga[kMomGroup], gb[kMomGroup];
  auto pass = [&](auto &&body) {
    mom_load_group(d, 0, lane, ga);
    for (int base = 0; base < n_max; base += 2 * kMomSpan) {
      mom_load_group(d, base + kMomSpan, lane, gb);
      body(base, ga);
      __builtin_amdgcn_sched_barrier(0);
      mom_load_group(d, base + 2 * kMomSpan, lane, ga);
      body(base + kMomSpan, gb);
      __builtin_amdgcn_sched_barrier(0);
    }
  };

  ExactSum ex, el;
  pass([&](int base, const MomLoad (&cur)[kMomGroup]) {
#pragma unroll
    for (int g = 0; g < kMomGroup; ++g) {
      const int i = base + 64 * g + lane;
      if (i < n_e) exact_add(ex, cur[g].x);
      if (i < nk) exact_add(el, levels[rank_lohi(cur[g].lo, cur[g].hi4 >> 16)]);
    }
  });
  double ev_sum, km_sum;
  const bool ex_ok = exact_total(ex, ev_sum), el_ok = exact_total(el, km_sum);
  if (!ex_ok || !el_ok) {
    double acc = 0.0;
    pass([&](int base, const MomLoad (&cur)[kMomGroup]) {
#pragma unroll
      for (int g = 0; g < kMomGroup; ++g) {
        const int i = base + 64 * g + lane;
        st[64 * g + lane] = i < n_e ? (double)cur[g].x : 0.0;
        st[kMomSecond + 64 * g + lane] =
            i < nk ? (double)levels[rank_lohi(cur[g].lo, cur[g].hi4 >> 16)] : 0.0;
      }
      wave_lds_sync();
      serial_sum_span(st + ((lane & 1) ? kMomSecond : 0), acc);
      wave_lds_sync();
    });
    if (!ex_ok) ev_sum = readlane_f64(acc, 0);
    if (!el_ok) km_sum = readlane_f64(acc, 1);
  }
  const double shift = ev_sum / n_e - km_sum / nk;

  double acc = 0.0;
  pass([&](int base, const MomLoad (&cur)[kMomGroup]) {
#pragma unroll
    for (int g = 0; g < kMomGroup; ++g) {
      const int i = base + 64 * g + lane;
      double l2 = 0.0, d2 = 0.0;
      if (i < nk) {
        const double l = levels[rank_lohi(cur[g].lo, cur[g].hi4 >> 16)];
        l2 = l * l;
      }
      if (i < n_e) {
        const double dx = (double)cur[g].x - shift;
        d2 = dx * dx;
      }
      st[64 * g + lane] = l2;
      st[kMomSecond + 64 * g + lane] = d2;
    }
    wave_lds_sync();
    serial_sum_span(st + ((lane & 1) ? kMomSecond : 0), acc);
    wave_lds_sync();
  });
  const double km_sq = readlane_f64(acc, 0), ev_sq = readlane_f64(acc, 1);
  const double scale = (ev_sq / n_e) / (km_sq / nk);
  if (lane == 0) {
    abea_scalings_t o;
    o.scale = (float)scale;
    o.shift = (float)shift;
    o.var = 0.0f;
    o.log_var = 0.0f;
    a.out[r] = o;
  }
}

}  // namespace

hipError_t launch_postalign(const PostArgs &a, hipStream_t s) {
  if (a.n_reads <= 0) return hipSuccess;
  if (a.n_items > 0) {
    const int wb = (a.n_items + kWalkWaves - 1) / kWalkWaves;
    auto walk = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(wb), dim3(64 * kWalkWaves), 0, s, a);
    };
    switch ((a.rec ? 4 : 0) | (a.bmap ? 2 : 0) | (a.recalibrate ? 1 : 0)) {
      case 0: walk(abea_postalign_walk_kernel<false, false, false>); break;
      case 1: walk(abea_postalign_walk_kernel<false, false, true>); break;
      case 2: walk(abea_postalign_walk_kernel<false, true, false>); break;
      case 3: walk(abea_postalign_walk_kernel<false, true, true>); break;
      case 4: walk(abea_postalign_walk_kernel<true, false, false>); break;
      case 5: walk(abea_postalign_walk_kernel<true, false, true>); break;
      case 6: walk(abea_postalign_walk_kernel<true, true, false>); break;
      default: walk(abea_postalign_walk_kernel<true, true, true>); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int blocks = (a.n_reads + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(abea_postalign_finish_kernel, dim3(blocks), dim3(64 * kWavesPerBlock), 0, s,
                     a);
  return hipGetLastError();
}

int postalign_segment_pairs() { return kSegPairs; }

hipError_t launch_recalibrate_records(const RecalArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(abea_recal_records_kernel, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mom(const MomArgs &a, hipStream_t s) {
  if (a.n_reads <= 0) return hipSuccess;
  const int blocks = (a.n_reads + kMomWaves - 1) / kMomWaves;
  hipLaunchKernelGGL(abea_mom_kernel, dim3(blocks), dim3(64 * kMomWaves), 0, s, a);
  return hipGetLastError();
}

}  // namespace abea
