"""abea_align_noalloc(): the per-read align() that allocates nothing (its
buffers come from a static device array).  Called through the C ABI, its
pairs must equal abea_align()'s (and f5c's golden pairs) bit for bit, call
after call, with a smaller read following a larger one (the array keeps the
previous call's contents); reads beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def _call(L, fn, seq, ev_aos, model, scale, shift):
    ev = np.ascontiguousarray(ev_aos)
    et = _table(ev)
    sc = _scal(scale, shift)
    out = np.zeros((2 * max(ev.shape[0], 1), 2), np.int32)
    n = fn(out.ctypes.data_as(C.c_void_p), seq, len(seq), et, model.ctypes.data_as(C.c_void_p),
           sc, 4000.0)
    return n, out[:max(n, 0)].copy()


def _table(ev):
    from abea_amd import _lib
    return _lib.EventTable(ev.shape[0], 0, ev.shape[0], ev.ctypes.data_as(C.POINTER(_lib.EventT)))


def _scal(scale, shift):
    from abea_amd import _lib
    return _lib.ScalingsT(float(scale), float(shift), 0.0, 0.0)


def _aos(means):
    from abea_amd.dump import EVENT_DTYPE
    a = np.zeros(len(means), EVENT_DTYPE)
    a["mean"] = means
    return a


def test_noalloc_equals_align_call_after_call():
    _need_gpu()
    from abea_amd import _lib, load_golden_subset, synth
    L = _lib.lib()
    b, golden, _ = load_golden_subset(GOLDEN)
    model = np.ascontiguousarray(b.model, np.float32)
    big = synth.generate(np.array([120000, 30000]), seed=77, model=b.model)
    cases = [(big.read_seq(0), _aos(big.read_events(0)), big.scale[0], big.shift[0], None)]
    aos = b.events_aos()
    order = np.argsort(-b.read_len)
    for i in order:  # each golden read, longest first: every call smaller than the last
        cases.append((b.read_seq(i), aos[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]], b.scale[i],
                      b.shift[i], golden[i]))
    cases.append((big.read_seq(1), _aos(big.read_events(1)), big.scale[1], big.shift[1], None))
    rng = np.random.default_rng(5)
    for nb, ne in [(5, 10), (6, 3), (400, 30), (60, 400), (50, 0)]:
        s = rng.choice(np.frombuffer(b"ACGTN", np.uint8), nb).tobytes()
        cases.append((s, _aos(rng.uniform(55, 135, ne).astype(np.float32)), 1.0, 0.0, None))
    for j, (seq, ev, scale, shift, gold) in enumerate(cases):
        print(f"[noalloc] case {j}: {len(seq)} bases, {len(ev)} events", flush=True)
        n0, p0 = _call(L, L.abea_align, seq, ev, model, scale, shift)
        n1, p1 = _call(L, L.abea_align_noalloc, seq, ev, model, scale, shift)
        assert n0 >= 0, L.abea_last_error()
        assert n1 == n0, (j, n1, n0, L.abea_last_error())
        assert np.array_equal(p1, p0), j
        if gold is not None:
            assert n1 == gold.shape[0] and np.array_equal(p1, gold), j
    # the same large read again after the small ones
    seq, ev, scale, shift, _ = cases[0]
    n0, p0 = _call(L, L.abea_align, seq, ev, model, scale, shift)
    n1, p1 = _call(L, L.abea_align_noalloc, seq, ev, model, scale, shift)
    assert n1 == n0 and np.array_equal(p1, p0)


def test_noalloc_refuses_reads_beyond_its_array():
    """A read whose buffers exceed the static array is refused on the host
    with a message naming the limit (nothing runs on the device)."""
    _need_gpu()
    from abea_amd import _lib, load_golden_subset
    L = _lib.lib()
    b, _, _ = load_golden_subset(GOLDEN)
    model = np.ascontiguousarray(b.model, np.float32)
    rng = np.random.default_rng(6)
    seq = rng.choice(np.frombuffer(b"ACGT", np.uint8), 400000).tobytes()
    ev = _aos(rng.uniform(60, 130, 660000).astype(np.float32))
    n, _ = _call(L, L.abea_align_noalloc, seq, ev, model, 1.0, 0.0)
    assert n == -1
    msg = L.abea_last_error().decode()
    assert "static array holds" in msg and "67108864" in msg, msg
