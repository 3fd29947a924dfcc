"""f5c dump I/O (GPU/dump_functions.h layout): write / read round trips on
the CPU; the GPU replay of a dump lives in test_gpu_parity.py."""
import os

import numpy as np

from conftest import GOLDEN, REFERENCE_DUMP


def _golden_dump_batch():
    from abea_amd import dump, load_golden_subset
    b, gold, _ = load_golden_subset(GOLDEN)
    return b, gold, dump.from_engine_batch(b, pairs=gold)


def test_dump_round_trip(tmp_path):
    from abea_amd import dump
    b, gold, db = _golden_dump_batch()
    d = str(tmp_path / "dump")
    dump.write_dump(d, [db, db])
    back = dump.load_dump(d)
    assert len(back) == 2
    for bb in back:
        assert np.array_equal(bb.model, db.model)
        assert len(bb.reads) == len(db.reads)
        for x, y in zip(bb.reads, db.reads):
            assert x.seq == y.seq
            assert x.events.tobytes() == y.events.tobytes()
            assert x.scalings.tobytes() == y.scalings.tobytes()
            assert x.sample_rate == y.sample_rate
            assert np.array_equal(x.pairs, y.pairs)
    eb = dump.to_engine_batch(back[0])
    assert np.array_equal(eb.seq, b.seq) and np.array_equal(eb.ev_mean, b.ev_mean)
    assert np.array_equal(eb.scalings(), b.scalings())
    assert eb.events_aos().tobytes() == b.events_aos().tobytes()


def test_reader_matches_reference_dump_files():
    """The reader on the reference's own dump directory (when present in this
    container) gives, read for read, the inputs and outputs of the committed
    golden subset, and write_batch reproduces those files byte for byte."""
    from abea_amd import dump
    if not os.path.isdir(REFERENCE_DUMP):
        import pytest
        pytest.skip("reference dump not mounted")
    from abea_amd import load_golden_subset
    b, gold, z = load_golden_subset(GOLDEN)
    idx = [int(i) for i in z["read_index"]]
    ref = dump.load_batch(os.path.join(REFERENCE_DUMP, "0"), limit=max(idx) + 1)
    assert np.array_equal(ref.model, b.model)
    for j, i in enumerate(idx):
        x = ref.reads[i]
        assert x.seq == b.read_seq(j)
        ev = b.events_aos()[int(b.ev_ptr[j]):int(b.ev_ptr[j]) + int(b.n_events[j])]
        for f in ("start", "length", "mean", "stdv"):
            assert np.array_equal(x.events[f], ev[f]), (i, f)
        assert np.array_equal(np.array(tuple(x.scalings), np.float32), b.scalings()[j])
        assert np.float32(x.sample_rate) == z["sample_rate"][j]
        assert np.array_equal(x.pairs, gold[j])
    # the writer reproduces the reference's own files (the event_table
    # header's pointer word aside)
    import tempfile
    i0 = idx[0]
    with tempfile.TemporaryDirectory() as td:
        one = dump.DumpBatch(model=ref.model, reads=[ref.reads[i0]])
        dump.write_batch(td, one)
        for name in ("read_len[i].dat", "read[i].dat", "scalings[i].dat",
                     "f5[i].sample_rate.dat", "n_event_align_pairs[i].dat",
                     "event_align_pairs.dat"):
            a = open(os.path.join(REFERENCE_DUMP, "0", str(i0), name), "rb").read()
            w = open(os.path.join(td, "0", name), "rb").read()
            assert a == w, name
        a = open(os.path.join(REFERENCE_DUMP, "0", str(i0), "et[i].dat"), "rb").read()
        w = open(os.path.join(td, "0", "et[i].dat"), "rb").read()
        assert len(a) == len(w) and a[:24] == w[:24] and a[32:] == w[32:]
        assert open(os.path.join(REFERENCE_DUMP, "0", "model.dat"), "rb").read() == \
            open(os.path.join(td, "model.dat"), "rb").read()
