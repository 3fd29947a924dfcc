"""Reader/writer for f5c's per-batch dump directories.

Format (GPU/dump_functions.h:1-149):
  <dir>/no_of_batches.dat             int32
  <dir>/<b>/n_bam_rec.dat             int32
  <dir>/<b>/model.dat                 4096 x model_t {f32 mean, stdv, log_stdv}
  <dir>/<b>/<i>/read_len[i].dat       int32
  <dir>/<b>/<i>/read[i].dat           read_len bytes (no NUL)
  <dir>/<b>/<i>/et[i].dat             event_table header (32 B: size_t n, start,
                                      end, event_t*) + n x event_t (24 B:
                                      u64 start, f32 length, mean, stdv, pad)
  <dir>/<b>/<i>/scalings[i].dat       scalings_t {f32 scale, shift, var, log_var}
  <dir>/<b>/<i>/f5[i].sample_rate.dat f32
  <dir>/<b>/<i>/n_event_align_pairs[i].dat  int32      (outputs)
  <dir>/<b>/<i>/event_align_pairs.dat n x {i32 ref_pos, read_pos}
Struct sizes are the x86-64 layouts of GPU/f5c.h:112-181.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

EVENT_DTYPE = np.dtype([("start", "<u8"), ("length", "<f4"), ("mean", "<f4"), ("stdv", "<f4"),
                        ("_pad", "<u4")])
assert EVENT_DTYPE.itemsize == 24
EVENT_TABLE_HEADER = 32
MODEL_DTYPE = np.dtype([("level_mean", "<f4"), ("level_stdv", "<f4"), ("level_log_stdv", "<f4")])
SCALINGS_DTYPE = np.dtype([("scale", "<f4"), ("shift", "<f4"), ("var", "<f4"), ("log_var", "<f4")])


@dataclass
class DumpRead:
    seq: bytes
    events: np.ndarray          # EVENT_DTYPE[n]
    scalings: np.ndarray        # SCALINGS_DTYPE scalar record
    sample_rate: float
    pairs: np.ndarray | None = None   # int32[n, 2] {ref_pos, read_pos}


@dataclass
class DumpBatch:
    model: np.ndarray           # float32[4096, 3]
    reads: list = field(default_factory=list)


def _read_i32(path):
    return int(np.fromfile(path, "<i4", 1)[0])


def load_batch(batch_dir: str, with_outputs: bool = True, limit: int | None = None) -> DumpBatch:
    n = _read_i32(os.path.join(batch_dir, "n_bam_rec.dat"))
    model = np.fromfile(os.path.join(batch_dir, "model.dat"), "<f4").reshape(4096, 3)
    batch = DumpBatch(model=model)
    for i in range(n if limit is None else min(n, limit)):
        d = os.path.join(batch_dir, str(i))
        L = _read_i32(os.path.join(d, "read_len[i].dat"))
        with open(os.path.join(d, "read[i].dat"), "rb") as f:
            seq = f.read(L)
        assert len(seq) == L
        raw = np.fromfile(os.path.join(d, "et[i].dat"), np.uint8)
        n_ev = int(raw[:8].view("<u8")[0])
        ev = raw[EVENT_TABLE_HEADER:EVENT_TABLE_HEADER + 24 * n_ev].view(EVENT_DTYPE)
        assert ev.shape[0] == n_ev
        sc = np.fromfile(os.path.join(d, "scalings[i].dat"), SCALINGS_DTYPE, 1)[0]
        sr = float(np.fromfile(os.path.join(d, "f5[i].sample_rate.dat"), "<f4", 1)[0])
        pairs = None
        if with_outputs:
            npairs = _read_i32(os.path.join(d, "n_event_align_pairs[i].dat"))
            pairs = np.fromfile(os.path.join(d, "event_align_pairs.dat"), "<i4",
                                2 * npairs).reshape(npairs, 2)
        batch.reads.append(DumpRead(seq=seq, events=ev.copy(), scalings=sc, sample_rate=sr,
                                    pairs=pairs))
    return batch


def load_dump(dump_dir: str, with_outputs: bool = True):
    nb = _read_i32(os.path.join(dump_dir, "no_of_batches.dat"))
    return [load_batch(os.path.join(dump_dir, str(b)), with_outputs) for b in range(nb)]


def write_outputs(batch_dir: str, n_pairs, pairs_list) -> None:
    """Write n_event_align_pairs[i].dat / event_align_pairs.dat per read in
    the same layout f5c's dump writer uses."""
    for i, (n, p) in enumerate(zip(n_pairs, pairs_list)):
        d = os.path.join(batch_dir, str(i))
        os.makedirs(d, exist_ok=True)
        np.asarray([n], "<i4").tofile(os.path.join(d, "n_event_align_pairs[i].dat"))
        np.ascontiguousarray(p[:n], "<i4").tofile(os.path.join(d, "event_align_pairs.dat"))


def write_batch(batch_dir: str, batch: DumpBatch, with_outputs: bool = True) -> None:
    """Write one batch directory in f5c's dump layout (inputs, and the
    outputs of reads that carry pairs).  The event_table header's pointer
    field is written as 0."""
    os.makedirs(batch_dir, exist_ok=True)
    np.asarray([len(batch.reads)], "<i4").tofile(os.path.join(batch_dir, "n_bam_rec.dat"))
    np.ascontiguousarray(batch.model, "<f4").tofile(os.path.join(batch_dir, "model.dat"))
    for i, rd in enumerate(batch.reads):
        d = os.path.join(batch_dir, str(i))
        os.makedirs(d, exist_ok=True)
        np.asarray([len(rd.seq)], "<i4").tofile(os.path.join(d, "read_len[i].dat"))
        with open(os.path.join(d, "read[i].dat"), "wb") as f:
            f.write(rd.seq)
        n = rd.events.shape[0]
        hdr = np.asarray([n, 0, n, 0], "<u8")  # n, start, end, event pointer
        with open(os.path.join(d, "et[i].dat"), "wb") as f:
            f.write(hdr.tobytes())
            f.write(np.ascontiguousarray(rd.events, EVENT_DTYPE).tobytes())
        np.asarray(rd.scalings, SCALINGS_DTYPE).tofile(os.path.join(d, "scalings[i].dat"))
        np.asarray([rd.sample_rate], "<f4").tofile(os.path.join(d, "f5[i].sample_rate.dat"))
        if with_outputs and rd.pairs is not None:
            np.asarray([rd.pairs.shape[0]], "<i4").tofile(
                os.path.join(d, "n_event_align_pairs[i].dat"))
            np.ascontiguousarray(rd.pairs, "<i4").tofile(os.path.join(d, "event_align_pairs.dat"))


def write_dump(dump_dir: str, batches, with_outputs: bool = True) -> None:
    """Write a dump directory: no_of_batches.dat and one directory per batch."""
    os.makedirs(dump_dir, exist_ok=True)
    np.asarray([len(batches)], "<i4").tofile(os.path.join(dump_dir, "no_of_batches.dat"))
    for b, batch in enumerate(batches):
        write_batch(os.path.join(dump_dir, str(b)), batch, with_outputs)


def to_engine_batch(batch: DumpBatch):
    """The dump batch as an abea_amd.synth.Batch (CSR arrays, event records
    kept for the AoS upload)."""
    from .synth import Batch
    n = len(batch.reads)
    rl = np.array([len(r.seq) for r in batch.reads], np.int32)
    ne = np.array([r.events.shape[0] for r in batch.reads], np.int32)
    sp = np.zeros(n, np.int64)
    ep = np.zeros(n, np.int64)
    if n > 1:
        sp[1:] = np.cumsum(rl.astype(np.int64))[:-1]
        ep[1:] = np.cumsum(ne.astype(np.int64))[:-1]
    ev = (np.concatenate([r.events for r in batch.reads]) if n else np.zeros(0, EVENT_DTYPE))
    sc = np.array([tuple(r.scalings) for r in batch.reads], np.float32).reshape(n, 4)
    seq = np.frombuffer(b"".join(r.seq for r in batch.reads), np.uint8).copy()
    return Batch(model=np.ascontiguousarray(batch.model, np.float32), seq=seq, seq_ptr=sp,
                 read_len=rl, ev_mean=np.ascontiguousarray(ev["mean"], np.float32), ev_ptr=ep,
                 n_events=ne, scale=sc[:, 0].copy(), shift=sc[:, 1].copy(), var=sc[:, 2].copy(),
                 log_var=sc[:, 3].copy(), ev_stdv=np.ascontiguousarray(ev["stdv"]),
                 ev_length=np.ascontiguousarray(ev["length"]),
                 ev_start=np.ascontiguousarray(ev["start"]))


def from_engine_batch(b, sample_rate: float = 4000.0, pairs=None) -> DumpBatch:
    """An abea_amd.synth.Batch as a dump batch (optionally with outputs)."""
    ev_all = b.events_aos()
    out = DumpBatch(model=np.ascontiguousarray(b.model, np.float32))
    sc = b.scalings()
    for i in range(b.n_reads):
        p = int(b.ev_ptr[i])
        ev = ev_all[p:p + int(b.n_events[i])].copy()
        rec = np.zeros((), SCALINGS_DTYPE)
        rec["scale"], rec["shift"], rec["var"], rec["log_var"] = sc[i]
        out.reads.append(DumpRead(seq=b.read_seq(i), events=ev, scalings=rec,
                                  sample_rate=float(sample_rate),
                                  pairs=None if pairs is None else np.asarray(pairs[i], np.int32)))
    return out
