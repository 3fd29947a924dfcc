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
