"""Seeded synthetic nanopore-like ABEA inputs.

Generative model (all float32, seeded numpy PCG64):
  pore model   4096 6-mers, level_mean ~ U(60,130) pA, level_stdv ~ U(1.5,3.5),
               level_log_stdv = log(level_stdv) computed once and stored (the
               reference reads the cached value, GPU/align_cuda.cl:517-518)
  read         uniform ACGT
  events       per k-mer: 0 events with p=0.1, else Geometric(p=0.55) on {1,2,..}
  event mean   scale*level_mean + shift + N(0, (var*level_stdv)^2)
  event stdv   U(1,3); length = U(0.002,0.01) s * 4000 Hz; start cumulative
  scalings     scale ~ N(1,0.05), shift ~ N(0,3), var ~ U(1,1.5), log_var = log var

A Batch is CSR-flattened like the reference's batch flattening
(GPU/align_cuda.c:423-492): one byte array of bases with int64 offsets, one
event array with int64 offsets.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SAMPLE_RATE = 4000.0
BASES = np.frombuffer(b"ACGT", np.uint8)


@dataclass
class Batch:
    model: np.ndarray       # float32[4096, 3]
    seq: np.ndarray         # uint8[sum L]
    seq_ptr: np.ndarray     # int64[n]
    read_len: np.ndarray    # int32[n]
    ev_mean: np.ndarray     # float32[sum n_e]
    ev_ptr: np.ndarray      # int64[n]
    n_events: np.ndarray    # int32[n]
    scale: np.ndarray       # float32[n]
    shift: np.ndarray
    var: np.ndarray
    log_var: np.ndarray
    ev_stdv: np.ndarray | None = None
    ev_length: np.ndarray | None = None
    ev_start: np.ndarray | None = None

    @property
    def n_reads(self) -> int:
        return int(self.read_len.shape[0])

    @property
    def n_kmers(self) -> np.ndarray:
        return np.maximum(self.read_len.astype(np.int64) - 5, 0)

    @property
    def n_bands(self) -> np.ndarray:
        return self.n_events.astype(np.int64) + self.n_kmers + 2

    def nominal_cells(self) -> int:
        """W * sum(n_events + n_kmers + 2), W = 100."""
        return int(100 * self.n_bands.sum())

    def read_seq(self, i: int) -> bytes:
        p = int(self.seq_ptr[i])
        return self.seq[p:p + int(self.read_len[i])].tobytes()

    def read_events(self, i: int) -> np.ndarray:
        p = int(self.ev_ptr[i])
        return self.ev_mean[p:p + int(self.n_events[i])]

    def scalings(self) -> np.ndarray:
        return np.stack([self.scale, self.shift, self.var, self.log_var], 1).astype(np.float32)

    def events_aos(self) -> np.ndarray:
        """event_t records (24 B, GPU/f5c.h:113-121) for the whole batch."""
        from .dump import EVENT_DTYPE
        n = self.ev_mean.shape[0]
        a = np.zeros(n, EVENT_DTYPE)
        a["mean"] = self.ev_mean
        if self.ev_stdv is not None:
            a["stdv"] = self.ev_stdv
            a["length"] = self.ev_length
            a["start"] = self.ev_start
        return a

    def subset(self, idx) -> "Batch":
        idx = np.asarray(idx, np.int64)
        seqs = [self.seq[self.seq_ptr[i]:self.seq_ptr[i] + self.read_len[i]] for i in idx]
        evs = [self.ev_mean[self.ev_ptr[i]:self.ev_ptr[i] + self.n_events[i]] for i in idx]
        rl = self.read_len[idx].astype(np.int32)
        ne = self.n_events[idx].astype(np.int32)
        sp = np.zeros(len(idx), np.int64)
        sp[1:] = np.cumsum(rl.astype(np.int64))[:-1]
        ep = np.zeros(len(idx), np.int64)
        ep[1:] = np.cumsum(ne.astype(np.int64))[:-1]
        return Batch(self.model, np.concatenate(seqs) if len(seqs) else np.zeros(0, np.uint8), sp,
                     rl, np.concatenate(evs) if len(evs) else np.zeros(0, np.float32), ep, ne,
                     self.scale[idx].copy(), self.shift[idx].copy(), self.var[idx].copy(),
                     self.log_var[idx].copy())


def make_model(rng: np.random.Generator) -> np.ndarray:
    mu = rng.uniform(60.0, 130.0, 4096).astype(np.float32)
    sd = rng.uniform(1.5, 3.5, 4096).astype(np.float32)
    return np.stack([mu, sd, np.log(sd)], 1).astype(np.float32)


def kmer_ranks(codes: np.ndarray) -> np.ndarray:
    """Rank of every 6-mer start in a 2-bit code array (codes[i] in 0..3),
    first base most significant (GPU/align_cuda.cl:39-51)."""
    n = codes.shape[0] - 5
    if n <= 0:
        return np.zeros(0, np.int32)
    r = np.zeros(n, np.int32)
    for i in range(6):
        r = (r << 2) | codes[i:i + n].astype(np.int32)
    return r


def generate(lengths, seed: int = 0, model: np.ndarray | None = None, with_aos: bool = False,
             chunk_bases: int = 1 << 24, skip: float = 0.1) -> Batch:
    """Generate reads of the given lengths (bases).

    Each k-mer emits a geometric(0.55) number of events (mean 1.8), or none
    with probability `skip` (10% in the benchmark workloads).  At 10% most
    reads fail f5c's QC on the average log-emission (each skipped k-mer adds
    a poorly matching cell to the path); skip = 0.015 gives about 0.9% skip
    pairs on the traced paths, the rate of the real reads in f5c's
    dump_small, which all pass QC."""
    lengths = np.asarray(lengths, np.int64)
    rng = np.random.default_rng(seed)
    if model is None:
        model = make_model(rng)
    n = lengths.shape[0]
    scale = rng.normal(1.0, 0.05, n).astype(np.float32)
    shift = rng.normal(0.0, 3.0, n).astype(np.float32)
    var = rng.uniform(1.0, 1.5, n).astype(np.float32)
    log_var = np.log(var).astype(np.float32)

    seq = BASES[rng.integers(0, 4, int(lengths.sum()), dtype=np.uint8)]
    seq_ptr = np.zeros(n, np.int64)
    seq_ptr[1:] = np.cumsum(lengths)[:-1]
    lut = np.zeros(256, np.uint8)
    lut[BASES] = np.arange(4, dtype=np.uint8)

    # process reads in chunks of bases to bound temporary memory
    ev_parts, ne_parts, sd_parts, len_parts = [], [], [], []
    start = 0
    while start < n:
        end = start
        acc = 0
        while end < n and (acc == 0 or acc + lengths[end] <= chunk_bases):
            acc += int(lengths[end])
            end += 1
        for i in range(start, end):
            pass
        s0 = int(seq_ptr[start])
        s1 = int(seq_ptr[end - 1] + lengths[end - 1])
        codes = lut[seq[s0:s1]]
        ranks_all = kmer_ranks(codes)
        # per read k-mer index ranges inside this chunk
        loc = seq_ptr[start:end] - s0
        nk = np.maximum(lengths[start:end] - 5, 0)
        sel = np.concatenate([np.arange(a, a + b) for a, b in zip(loc, nk)]) if len(nk) else \
            np.zeros(0, np.int64)
        ranks = ranks_all[sel]
        nkt = ranks.shape[0]
        cnt = rng.geometric(0.55, nkt).astype(np.int32)
        cnt[rng.random(nkt) < skip] = 0
        read_of_kmer = np.repeat(np.arange(start, end), nk)
        ev_rank = np.repeat(ranks, cnt)
        ev_read = np.repeat(read_of_kmer, cnt)
        mu = model[ev_rank, 0]
        sd = model[ev_rank, 1]
        noise = rng.standard_normal(ev_rank.shape[0]).astype(np.float32)
        mean = (scale[ev_read] * mu + shift[ev_read] +
                noise * (var[ev_read] * sd)).astype(np.float32)
        ev_parts.append(mean)
        kc = np.zeros(nkt + 1, np.int64)
        np.cumsum(cnt, out=kc[1:])
        bounds = np.concatenate([[0], np.cumsum(nk)])
        ne_parts.append((kc[bounds[1:]] - kc[bounds[:-1]]).astype(np.int32))
        if with_aos:
            sd_parts.append(rng.uniform(1.0, 3.0, mean.shape[0]).astype(np.float32))
            len_parts.append(np.floor(rng.uniform(0.002, 0.01, mean.shape[0]) *
                                      SAMPLE_RATE).astype(np.float32))
        start = end
    ev_mean = np.concatenate(ev_parts) if ev_parts else np.zeros(0, np.float32)
    n_events = np.concatenate(ne_parts) if ne_parts else np.zeros(0, np.int32)
    ev_ptr = np.zeros(n, np.int64)
    ev_ptr[1:] = np.cumsum(n_events.astype(np.int64))[:-1]
    b = Batch(model, seq, seq_ptr, lengths.astype(np.int32), ev_mean, ev_ptr, n_events, scale,
              shift, var, log_var)
    if with_aos:
        b.ev_stdv = np.concatenate(sd_parts)
        b.ev_length = np.concatenate(len_parts)
        st = np.zeros(ev_mean.shape[0], np.uint64)
        st[1:] = np.cumsum(b.ev_length[:-1].astype(np.uint64))
        b.ev_start = st
    return b


def lognormal_lengths(n: int, median: float, sigma: float, lo: int, hi: int, seed: int):
    rng = np.random.default_rng(seed + 7919)
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, n)), lo, hi).astype(np.int64)


# Workloads named by what they are (reads x length distribution).
WORKLOADS = {
    "cpu64x2kb": dict(n=64, fixed=2000),
    "plain4096_ln5kb": dict(n=4096, median=5000, sigma=0.5, lo=500, hi=50000),
    "batched16384_ln8kb": dict(n=16384, median=8000, sigma=0.52, lo=500, hi=200000),
    "adjoint1024x10kb": dict(n=1024, fixed=10000),
    "refine8192x20kb": dict(n=8192, fixed=20000),
    # reads that pass QC (real-data skip rate), for postalign/recalibration
    "postalign16384_ln8kb": dict(n=16384, median=8000, sigma=0.52, lo=500, hi=200000,
                                 skip=0.015),
}


def workload_lengths(name: str, seed: int = 0, n: int | None = None) -> np.ndarray:
    w = WORKLOADS[name]
    cnt = w["n"] if n is None else n
    if "fixed" in w:
        return np.full(cnt, w["fixed"], np.int64)
    return lognormal_lengths(w["n"], w["median"], w["sigma"], w["lo"], w["hi"], seed)[:cnt]


def workload(name: str, seed: int = 0, n: int | None = None, with_aos: bool = False) -> Batch:
    return generate(workload_lengths(name, seed, n), seed=seed, with_aos=with_aos,
                    skip=WORKLOADS[name].get("skip", 0.1))


def raw_signal(b: Batch, seed: int = 0):
    """A raw signal (float32 pA) for the batch's events, the input of
    getevents: each event lasts floor(U(0.002, 0.01) s * 4000 Hz) samples
    (its recorded length when the batch has them) of its mean plus N(0, sd)
    noise, sd ~ U(1, 3) (its recorded stdv when present).  Returns (raw,
    raw_off) with read i at raw[raw_off[i]:raw_off[i+1]]."""
    rng = np.random.default_rng(seed + 0x5161)
    ne = b.ev_mean.shape[0]
    if b.ev_length is not None:
        lens = b.ev_length.astype(np.int64)
    else:
        lens = np.floor(rng.uniform(0.002, 0.01, ne) * SAMPLE_RATE).astype(np.int64)
    sd = b.ev_stdv if b.ev_stdv is not None else rng.uniform(1.0, 3.0, ne).astype(np.float32)
    raw = np.repeat(b.ev_mean, lens)
    raw += np.repeat(sd, lens) * rng.standard_normal(raw.shape[0], dtype=np.float32)
    per_read = np.zeros(b.n_reads, np.int64)
    read_of_event = np.repeat(np.arange(b.n_reads), b.n_events)
    np.add.at(per_read, read_of_event, lens)
    raw_off = np.zeros(b.n_reads + 1, np.int64)
    np.cumsum(per_read, out=raw_off[1:])
    return raw.astype(np.float32), raw_off


def perturb_scalings(b: Batch, seed: int = 0, shift_pm: float = 5.0, scale_pm: float = 0.10):
    """The refinement workload's start: each read's shift moved by
    U(-shift_pm, shift_pm) and its scale by a factor 1 + U(-scale_pm,
    scale_pm).  Modifies b in place; returns the generating (scale, shift)."""
    rng = np.random.default_rng(seed + 0x5CA1E)
    true_scale, true_shift = b.scale.copy(), b.shift.copy()
    b.scale = (b.scale * (1.0 + rng.uniform(-scale_pm, scale_pm, b.n_reads))).astype(np.float32)
    b.shift = (b.shift + rng.uniform(-shift_pm, shift_pm, b.n_reads)).astype(np.float32)
    return true_scale, true_shift
