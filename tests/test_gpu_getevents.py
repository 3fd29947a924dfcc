"""GPU parity of event detection (f5c getevents(), GPU/f5cmisc.h:44) with the
oracle restatement (oracle/getevents.c): event tables byte for byte.

Covers signal lengths around the window and chunk sizes, empty and tiny
signals, the serial prefix-sum path (terms too fine for an exact parallel
sum), exact signals too large or too fine for the statistics' fast
divisions, all-zero and subnormal signals, inf and NaN samples, forced
speculation misses (no lead-in), dyadic step signals with known cuts, reads
at a non-zero offset, and the host drop-ins."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "abea-opencl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from abea_amd import _lib, engine, synth  # noqa: E402


def _first_diff(a: np.ndarray, b: np.ndarray) -> str:
    if a.shape != b.shape:
        return f"shapes {a.shape} vs {b.shape}"
    ab, bb = a.view(np.uint8).reshape(a.shape[0], -1), b.view(np.uint8).reshape(b.shape[0], -1)
    bad = np.nonzero((ab != bb).any(1))[0]
    return "equal" if bad.size == 0 else f"first differing event {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}"


def _events_equal(got: np.ndarray, want: np.ndarray) -> bool:
    # compare the 20 meaningful bytes of each record (the 4 pad bytes are not
    # part of event_t's value)
    if got.shape != want.shape:
        return False
    g = got.view(np.uint8).reshape(-1, 24)[:, :20]
    w = want.view(np.uint8).reshape(-1, 24)[:, :20]
    return bool(np.array_equal(g, w))


def _signals(lengths, seed):
    """Slices of a synthetic raw signal with the given lengths."""
    b = synth.workload("plain4096_ln5kb", seed=seed, n=16)
    raw, off = synth.raw_signal(b, seed=seed)
    rng = np.random.default_rng(seed)
    out = []
    for L in lengths:
        s = int(rng.integers(0, raw.shape[0] - L)) if L < raw.shape[0] else 0
        out.append(raw[s:s + L].copy())
    return out


def _run_device(ctx, sigs, lead=0, overlap=None):
    lens = np.array([s.shape[0] for s in sigs], np.int64)
    raw_off = np.zeros(len(sigs) + 1, np.int64)
    raw_off[0] = lead
    raw_off[1:] = lead + np.cumsum(lens)
    flat = np.concatenate([np.full(lead, 77.0, np.float32)] + sigs) if sigs else np.zeros(1, np.float32)
    raw = torch.from_numpy(flat).cuda()
    if overlap is not None:
        _lib.check(_lib.lib().abea_ctx_set_getevents_overlap(ctx.handle, overlap), "overlap")
    ev, cnt = engine.getevents(ctx, raw, raw_off)
    torch.cuda.synchronize()
    res = engine.getevents_results(ev, cnt, raw_off)
    if overlap is not None:
        _lib.check(_lib.lib().abea_ctx_set_getevents_overlap(ctx.handle, 64), "overlap")
    return res


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


LENGTHS = [0, 1, 5, 6, 11, 12, 13, 100, 2047, 2048, 2049, 4200, 10000, 65536 + 17]


def test_lengths_against_oracle(ctx):
    sigs = _signals(LENGTHS, seed=3)
    got = _run_device(ctx, sigs)
    for L, s, g in zip(LENGTHS, sigs, got):
        want = oracle.getevents(s)
        print(f"[getevents] length {L}: {want.shape[0]} events", flush=True)
        assert _events_equal(g, want), f"length {L}: {_first_diff(g, want)}"


def test_whole_reads_and_reruns(ctx):
    b = synth.workload("plain4096_ln5kb", seed=5, n=24)
    raw, off = synth.raw_signal(b, seed=5)
    sigs = [raw[off[i]:off[i + 1]] for i in range(b.n_reads)]
    want = [oracle.getevents(s) for s in sigs]
    for overlap in (64, 0, 17, 512):
        got = _run_device(ctx, sigs, overlap=overlap)
        reruns = _lib.lib().abea_ctx_getevents_reruns(ctx.handle)
        print(f"[getevents] overlap {overlap}: chunks re-run {reruns}", flush=True)
        for i, (g, w) in enumerate(zip(got, want)):
            assert _events_equal(g, w), f"overlap {overlap} read {i}: {_first_diff(g, w)}"
        if overlap == 0:
            assert reruns > 0  # the re-run path was exercised


def test_detector_chunk_sizes_agree(ctx):
    """The detector chunk size (chosen per call unless fixed) changes only
    which lane walks which samples: every size gives the oracle's events."""
    b = synth.workload("plain4096_ln5kb", seed=6, n=12)
    raw, off = synth.raw_signal(b, seed=6)
    sigs = [raw[off[i]:off[i + 1]] for i in range(b.n_reads)] + _signals([700, 3], seed=6)
    want = [oracle.getevents(s) for s in sigs]
    L = _lib.lib()
    for chunk in (256, 512, 1000, 4096, 0):
        _lib.check(L.abea_ctx_set_getevents_chunk(ctx.handle, chunk), "chunk")
        got = _run_device(ctx, sigs)
        used = L.abea_ctx_getevents_chunk(ctx.handle)
        print(f"[getevents] chunk {chunk}: used {used}", flush=True)
        assert used == chunk or chunk == 0
        for i, (g, w) in enumerate(zip(got, want)):
            assert _events_equal(g, w), f"chunk {chunk} read {i}: {_first_diff(g, w)}"
    assert L.abea_ctx_set_getevents_chunk(ctx.handle, 100) != 0


def test_serial_prefix_path(ctx):
    # terms as fine as 2^-60 cannot be summed exactly in double over a pA
    # signal: the prefix sums take the serial path
    sigs = _signals([3000, 5000], seed=7)
    for s in sigs:
        s[::97] = np.float32(1e-18)
    got = _run_device(ctx, sigs)
    for s, g in zip(sigs, got):
        assert _events_equal(g, oracle.getevents(s)), _first_diff(g, oracle.getevents(s))


def test_large_samples_take_ieee_statistics(ctx):
    # pA signals scaled by 2^54 (largest samples above 2^60, their squares
    # still finite) keep their sums exact but leave the tame range, so their
    # statistics divide by IEEE division.  A power of two scales every sum,
    # mean and square exactly, and these signals are cut where the unscaled
    # ones are
    sigs = _signals([5000, 777, 12000], seed=13)
    big = [(s * np.float32(2.0 ** 54)).astype(np.float32) for s in sigs]
    assert all(float(np.abs(b).max()) >= 2.0 ** 60 for b in big)
    got = _run_device(ctx, big)
    small = _run_device(ctx, sigs)
    for b, g, h in zip(big, got, small):
        want = oracle.getevents(b)
        assert _events_equal(g, want), _first_diff(g, want)
        assert np.array_equal(g["start"], h["start"]) and np.array_equal(g["length"], h["length"])


def test_step_signals_and_offset_reads(ctx):
    rng = np.random.default_rng(9)
    sigs, cuts = [], []
    for n_seg in (2, 40, 300):
        lengths = rng.integers(7, 41, n_seg)
        levels = [100.0]
        while len(levels) < n_seg:
            v = round(float(rng.uniform(60, 130)) * 4) / 4
            if abs(v - levels[-1]) >= 5:
                levels.append(v)
        sigs.append(np.concatenate([np.full(L, v, np.float32) for L, v in zip(lengths, levels)]))
        cuts.append(np.cumsum(lengths)[:-1])
    sigs.insert(1, np.zeros(0, np.float32))  # an empty read between others
    got = _run_device(ctx, sigs, lead=5)
    for s, g in zip(sigs, got):
        assert _events_equal(g, oracle.getevents(s))
    # every step is a boundary; the last boundary ends no event
    for c, g in zip(cuts, [got[0], got[2], got[3]]):
        assert np.array_equal(g["start"][1:].astype(np.int64), c[:-1])
        assert g.shape[0] == max(len(c), 1)


def test_tiny_scales_and_the_tame_boundary(ctx):
    """Signals scaled by powers of two on either side of the tame reads'
    limit (samples and float squares on a 2^-125 grid): x 2^-50 stays tame
    (squares near 2^-86), x 2^-62 and x 2^-80 are exact but not tame
    (squares below the grid; subnormal at 2^-80); an all-zero read and a
    zero read with a few subnormal samples are exact.  Each read equals the
    oracle; the tame one is cut where the unscaled signal is."""
    base = _signals([5000, 3000, 4000], seed=31)
    sigs = [(base[0] * np.float32(2.0 ** -50)).astype(np.float32),
            (base[1] * np.float32(2.0 ** -62)).astype(np.float32),
            (base[2] * np.float32(2.0 ** -80)).astype(np.float32),
            np.zeros(3000, np.float32), np.zeros(2500, np.float32)]
    sigs[4][::301] = np.float32(2.0 ** -149)
    got = _run_device(ctx, sigs)
    for i, (s, g) in enumerate(zip(sigs, got)):
        want = oracle.getevents(s)
        assert _events_equal(g, want), f"read {i}: {_first_diff(g, want)}"
    unscaled = _run_device(ctx, [base[0]])[0]
    assert np.array_equal(got[0]["start"], unscaled["start"])
    assert np.array_equal(got[0]["length"], unscaled["length"])


def test_non_finite_samples(ctx):
    """A read with an inf sample and one with a NaN sample fail the exactness
    test and take the serial prefix sums; their events equal the oracle's,
    field by field, a NaN matching a NaN of either sign (x86 and the GPU
    make NaNs of different signs)."""
    base = _signals([4000, 4000, 3000], seed=41)
    sigs = [base[0].copy(), base[1].copy(), base[2]]
    sigs[0][1500] = np.float32(np.inf)
    sigs[1][2500] = np.float32(np.nan)
    got = _run_device(ctx, sigs)
    for i, (s, g) in enumerate(zip(sigs, got)):
        want = oracle.getevents(s)
        assert g.shape == want.shape, f"read {i}: {g.shape} vs {want.shape}"
        assert np.array_equal(g["start"], want["start"]), f"read {i}"
        assert np.array_equal(g["length"].view(np.uint32), want["length"].view(np.uint32)), i
        for f in ("mean", "stdv"):
            a, b = g[f], want[f]
            same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
            assert same.all(), f"read {i} {f}: first at {np.nonzero(~same)[0][:3]}"
    assert np.isinf(got[0]["mean"]).sum() == 1 and np.isnan(got[1]["mean"]).sum() == 1


def test_mixed_kinds_in_one_call(ctx):
    """Reads of every kind next to each other in one call: a tame read, then
    one scaled by 2^54 (IEEE divisions in its statistics), then one with
    1e-18 samples (serial prefix sums), then a tame one.  Near a read's end a
    detector lane loads the next read's samples, which must enter no
    statistic of its own read; each read must equal the oracle on its own."""
    base = _signals([4100, 3000, 2600, 5000], seed=21)
    sigs = [base[0], (base[1] * np.float32(2.0 ** 54)).astype(np.float32), base[2].copy(),
            base[3]]
    sigs[2][::89] = np.float32(1e-18)
    got = _run_device(ctx, sigs)
    for i, (s, g) in enumerate(zip(sigs, got)):
        want = oracle.getevents(s)
        assert _events_equal(g, want), f"read {i}: {_first_diff(g, want)}"
    # and in the other order, the fine and the large reads first
    rev = sigs[::-1]
    got = _run_device(ctx, rev)
    for i, (s, g) in enumerate(zip(rev, got)):
        want = oracle.getevents(s)
        assert _events_equal(g, want), f"reversed read {i}: {_first_diff(g, want)}"


def test_host_dropins(ctx):
    sigs = _signals([9000, 0, 31, 4097], seed=11)
    L = _lib.lib()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    # per read
    for s in sigs:
        et = L.abea_getevents(s.shape[0], s.ctypes.data_as(C.c_void_p) if s.shape[0] else None)
        want = oracle.getevents(s)
        assert et.n == want.shape[0] and et.start == 0 and et.end == et.n
        if et.n:
            buf = C.string_at(C.cast(et.event, C.c_void_p), 24 * et.n)
            got = np.frombuffer(buf, engine.EVENT_DTYPE)
            assert _events_equal(got, want), _first_diff(got, want)
            libc.free(C.cast(et.event, C.c_void_p))
        else:
            assert not et.event
    # batch
    n = len(sigs)
    ns = (C.c_size_t * n)(*[s.shape[0] for s in sigs])
    ptrs = (C.c_void_p * n)(*[s.ctypes.data if s.shape[0] else None for s in sigs])
    ets = (_lib.EventTable * n)()
    _lib.check(L.abea_getevents_db(n, ns, ptrs, ets), "abea_getevents_db")
    for s, et in zip(sigs, ets):
        want = oracle.getevents(s)
        assert et.n == want.shape[0]
        if et.n:
            got = np.frombuffer(C.string_at(C.cast(et.event, C.c_void_p), 24 * et.n),
                                engine.EVENT_DTYPE)
            assert _events_equal(got, want)
            libc.free(C.cast(et.event, C.c_void_p))
    # a null signal with samples is an error: an empty table and a message
    et = L.abea_getevents(10, None)
    assert et.n == 0 and not et.event
    assert b"null" in L.abea_last_error()
