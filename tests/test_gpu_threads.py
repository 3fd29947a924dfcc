"""The host-buffer drop-ins under f5c's threading: f5c calls align() from a
fresh set of worker threads for every batch (GPU/f5c.h:57-59), one read per
call.  Calls from many threads, started and joined over several batches,
must each return the reference's pairs and hold a bounded number of device
contexts (the pool behind the host entry points)."""
import ctypes as C
import threading

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def test_align_from_threads_over_batches():
    _need_gpu()
    from abea_amd import _lib, load_golden_subset
    L = _lib.lib()
    b, golden, _ = load_golden_subset(GOLDEN)
    model = np.ascontiguousarray(b.model, np.float32)
    aos = b.events_aos()
    n = 6
    seqs = [b.read_seq(i) for i in range(n)]
    evs = [np.ascontiguousarray(aos[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]]) for i in range(n)]
    n_threads, rounds = 8, 4
    before = L.abea_host_contexts()
    errors = []

    def worker(t):
        try:
            for j in range(n):
                i = (t + j) % n
                ev = evs[i]
                et = _lib.EventTable(ev.shape[0], 0, ev.shape[0],
                                     ev.ctypes.data_as(C.POINTER(_lib.EventT)))
                sc = _lib.ScalingsT(float(b.scale[i]), float(b.shift[i]), 0.0, 0.0)
                out = np.zeros((2 * ev.shape[0], 2), np.int32)
                got = L.abea_align(out.ctypes.data_as(C.c_void_p), seqs[i], len(seqs[i]), et,
                                   model.ctypes.data_as(C.c_void_p), sc, 4000.0)
                if got != golden[i].shape[0] or not np.array_equal(out[:got], golden[i]):
                    errors.append((t, i, got, L.abea_last_error()))
        except Exception as e:  # noqa: BLE001
            errors.append((t, repr(e)))

    free = []
    for rnd in range(rounds):
        ts = [threading.Thread(target=worker, args=(t,)) for t in range(n_threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
        print(f"[threads] round {rnd}: contexts {L.abea_host_contexts()}, free {free[-1]}",
              flush=True)
    assert errors == []
    created = L.abea_host_contexts() - before
    assert 0 <= created <= n_threads, created
    # later batches reuse the pooled contexts: no per-thread growth
    assert free[-1] >= free[0] - (64 << 20), free


def test_gathered_calls_with_two_models():
    """Concurrent align() calls are gathered into batches of one model each:
    threads that pass two different model arrays (equal contents) all get
    the reference's pairs, and calls with a model whose contents differ get
    that model's results (the oracle's)."""
    _need_gpu()
    import oracle
    from abea_amd import _lib, load_golden_subset
    L = _lib.lib()
    b, golden, _ = load_golden_subset(GOLDEN)
    model_a = np.ascontiguousarray(b.model, np.float32)
    model_b = model_a.copy()
    # a third model with shifted levels: its reads align differently
    model_c = model_a.copy()
    model_c[:, 0] += 1.5
    aos = b.events_aos()
    n = 6
    seqs = [b.read_seq(i) for i in range(n)]
    evs = [np.ascontiguousarray(aos[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]]) for i in range(n)]
    want_c = []
    for i in range(n):
        r = oracle.align(seqs[i], b.ev_mean[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]], model_c,
                         float(b.scale[i]), float(b.shift[i]))
        want_c.append(r["pairs"])
    errors = []

    def worker(t):
        try:
            model = (model_a, model_b, model_c)[t % 3]
            for j in range(2 * n):
                i = (t + j) % n
                ev = evs[i]
                et = _lib.EventTable(ev.shape[0], 0, ev.shape[0],
                                     ev.ctypes.data_as(C.POINTER(_lib.EventT)))
                sc = _lib.ScalingsT(float(b.scale[i]), float(b.shift[i]), 0.0, 0.0)
                out = np.zeros((2 * ev.shape[0], 2), np.int32)
                got = L.abea_align(out.ctypes.data_as(C.c_void_p), seqs[i], len(seqs[i]), et,
                                   model.ctypes.data_as(C.c_void_p), sc, 4000.0)
                want = want_c[i] if t % 3 == 2 else golden[i]
                if got != want.shape[0] or not np.array_equal(out[:got], want):
                    errors.append((t, i, got, want.shape[0], L.abea_last_error()))
        except Exception as e:  # noqa: BLE001
            errors.append((t, repr(e)))

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(9)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert errors == [], errors[:4]
    assert any(not np.array_equal(want_c[i], golden[i]) for i in range(n))


def test_align_db_phase_times():
    """abea_host_last_times() reports the phases of the last align_cuda()
    drop-in call, as the reference's host prints them."""
    _need_gpu()
    from abea_amd import _lib, load_golden_subset
    L = _lib.lib()
    b, golden, _ = load_golden_subset(GOLDEN)
    model = np.ascontiguousarray(b.model, np.float32)
    aos = b.events_aos()
    n = b.n_reads
    keep = []
    reads, ets = (C.c_char_p * n)(), (_lib.EventTable * n)()
    for i in range(n):
        ev = np.ascontiguousarray(aos[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]])
        keep.append(ev)
        reads[i] = b.read_seq(i)
        ets[i] = _lib.EventTable(ev.shape[0], 0, ev.shape[0],
                                 ev.ctypes.data_as(C.POINTER(_lib.EventT)))
    rl = np.ascontiguousarray(b.read_len, np.int32)
    scal = np.ascontiguousarray(b.scalings(), np.float32)
    outp = (C.POINTER(_lib.AlignedPair) * n)()
    nout = np.zeros(n, np.int32)
    st = L.abea_align_db(n, C.cast(reads, C.c_void_p), rl.ctypes.data_as(C.c_void_p),
                         C.cast(ets, C.c_void_p), model.ctypes.data_as(C.c_void_p),
                         scal.ctypes.data_as(C.c_void_p), C.cast(outp, C.c_void_p),
                         nout.ctypes.data_as(C.c_void_p))
    assert st == 0, L.abea_last_error()
    t = (C.c_double * 7)()
    L.abea_host_last_times(t)
    t = list(t)
    assert all(x >= 0 for x in t) and t[3] > 0 and t[6] > 0, t
    assert t[6] >= t[0] + t[5], t
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free.restype = None
    for i in range(n):
        assert nout[i] == golden[i].shape[0]
        got = np.ctypeslib.as_array(C.cast(outp[i], C.POINTER(C.c_int32)), (int(nout[i]) * 2,))
        assert np.array_equal(got.reshape(-1, 2), golden[i])
        libc.free(C.cast(outp[i], C.c_void_p))
