"""Repack a subset of the reference's golden f5c dump into a small fixture.

Source: FPGA/code/bins/dump_small/ (1 batch, 143 reads; format per
GPU/dump_functions.h:1-149).  The fixture keeps, per read, the sequence, the
event means (the only event field ABEA reads), the full event records, the
scalings and f5c's event_align_pairs; plus the 4096-row pore model.

Run in a container where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference/FPGA/code/bins/dump_small
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "abea-opencl_amd"))
from abea_amd import dump  # noqa: E402

N_KEEP = 24


def main(dump_dir):
    b = dump.load_batch(os.path.join(dump_dir, "0"))
    n = len(b.reads)
    lens = np.array([len(r.seq) for r in b.reads])
    order = np.argsort(lens, kind="stable")
    # spread the kept reads over the length distribution, shortest and longest included
    pick = sorted(set(order[np.linspace(0, n - 1, N_KEEP).round().astype(int)].tolist()))
    seqs, ev_ptr, pair_ptr = [], [0], [0]
    means, events, pairs, scal, idx = [], [], [], [], []
    for i in pick:
        r = b.reads[i]
        seqs.append(np.frombuffer(r.seq, np.uint8))
        means.append(r.events["mean"].astype(np.float32))
        events.append(r.events)
        pairs.append(r.pairs.astype(np.int32))
        scal.append([r.scalings["scale"], r.scalings["shift"], r.scalings["var"],
                     r.scalings["log_var"]])
        ev_ptr.append(ev_ptr[-1] + len(r.events))
        pair_ptr.append(pair_ptr[-1] + len(r.pairs))
        idx.append(i)
    seq_len = np.array([len(s) for s in seqs], np.int32)
    seq_ptr = np.concatenate([[0], np.cumsum(seq_len)]).astype(np.int64)
    out = os.path.join(HERE, "dump_small_subset.npz")
    np.savez_compressed(
        out,
        read_index=np.array(idx, np.int32),
        seq=np.concatenate(seqs),
        seq_ptr=seq_ptr,
        seq_len=seq_len,
        ev_mean=np.concatenate(means),
        ev_start=np.concatenate([e["start"] for e in events]).astype(np.uint64),
        ev_length=np.concatenate([e["length"] for e in events]).astype(np.float32),
        ev_stdv=np.concatenate([e["stdv"] for e in events]).astype(np.float32),
        ev_ptr=np.array(ev_ptr, np.int64),
        scalings=np.array(scal, np.float32),
        pairs=np.concatenate(pairs),
        pair_ptr=np.array(pair_ptr, np.int64),
        model=b.model.astype(np.float32),
        sample_rate=np.array([r.sample_rate for r in b.reads], np.float32)[pick],
    )
    print(f"wrote {out}: {len(pick)} reads, {os.path.getsize(out) / 1e6:.2f} MB")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/FPGA/code/bins/dump_small")
