"""GPU parity of the post-align and MOM kernels against the oracle: f5c's
postalign + recalibrate_model + read status after align(), and
estimate_scalings_using_mom.  Integer outputs (records, map, counts, flags)
and the float scalings must match bit for bit; events_per_base too."""
import ctypes as C

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from abea_amd import engine
    c = engine.Context(0)
    yield c
    c.close()


def golden_batch():
    from abea_amd import load_golden_subset
    return load_golden_subset(GOLDEN)


def qc_mixed_batch(n=48, seed=21):
    """Reads that pass align()'s QC (the real-data skip rate), reads that fail
    it (10% k-mer skips), and short reads with fewer than 200 'M' records."""
    from abea_amd import synth
    half = n // 2
    good_len = np.concatenate([synth.workload_lengths("postalign16384_ln8kb", seed, half - 4),
                               np.array([150, 180, 330, 700])])
    a = synth.generate(good_len, seed=seed, skip=0.015)
    b = synth.generate(synth.workload_lengths("batched16384_ln8kb", seed + 1, n - half),
                       seed=seed + 1, skip=0.1, model=a.model)
    return _concat(_concat(a, b), stay_heavy_reads(a.model, seed + 2))


def stay_heavy_reads(model, seed, n=3, n_k=400, per_kmer=6):
    """Reads with `per_kmer` events on every k-mer, close to the model
    levels: align() and the calibration pass, and events_per_base ~ 6 fails
    f5c's stay filter (FAILED_QUALITY_CHK)."""
    from abea_amd import synth
    rng = np.random.default_rng(seed)
    seqs, evs = [], []
    for _ in range(n):
        codes = rng.integers(0, 4, n_k + 5)
        ranks = synth.kmer_ranks(codes)
        mu, sd = model[ranks, 0], model[ranks, 1]
        ev = np.repeat(mu, per_kmer) + rng.normal(0, 0.5, n_k * per_kmer) * np.repeat(sd, per_kmer)
        seqs.append(synth.BASES[codes].tobytes())
        evs.append(ev.astype(np.float32))
    rl = np.array([len(q) for q in seqs], np.int32)
    ne = np.array([len(e) for e in evs], np.int32)
    sp = np.concatenate([[0], np.cumsum(rl.astype(np.int64))[:-1]])
    ep = np.concatenate([[0], np.cumsum(ne.astype(np.int64))[:-1]])
    return synth.Batch(model, np.frombuffer(b"".join(seqs), np.uint8).copy(), sp, rl,
                       np.concatenate(evs), ep, ne, np.ones(n, np.float32),
                       np.zeros(n, np.float32), np.ones(n, np.float32), np.zeros(n, np.float32))


def _concat(a, b):
    from abea_amd import synth
    seq = np.concatenate([a.seq, b.seq])
    sp = np.concatenate([a.seq_ptr, b.seq_ptr + a.seq.shape[0]])
    ev = np.concatenate([a.ev_mean, b.ev_mean])
    ep = np.concatenate([a.ev_ptr, b.ev_ptr + a.ev_mean.shape[0]])
    cat = lambda f: np.concatenate([getattr(a, f), getattr(b, f)])  # noqa: E731
    return synth.Batch(a.model, seq, sp, cat("read_len"), ev, ep, cat("n_events"), cat("scale"),
                       cat("shift"), cat("var"), cat("log_var"))


def run_post(ctx, batch, form=1, scale_var=True, records=True, base_map=True, aos=False):
    from abea_amd import engine
    db = engine.DeviceBatch.upload(batch, 0, aos_events=aos)
    plan = engine.Plan(ctx, db, form)
    plan.forward()
    o = plan.postalign(scale_var=scale_var, records=records, base_map=base_map)
    res = engine.postalign_results(plan, o)
    res["n_pairs"] = plan.n_pairs[:batch.n_reads].cpu().numpy()
    plan.close()
    return res


def _rec_tuple(r):
    return (bytes(r["ref_kmer"]), int(r["ref_position"]), int(r["read_idx"]),
            int(r["event_idx"]), int(r["rc"]), bytes(r["model_kmer"]), bytes(r["hmm_state"]))


def compare_post(batch, res, scale_var=True, records=True, base_map=True):
    import oracle
    bad = []
    sc_all = batch.scalings()
    for i in range(batch.n_reads):
        o = oracle.process_read(batch.read_seq(i), batch.read_events(i), batch.model, sc_all[i],
                                scale_var)
        ok = int(res["read_stat_flag"][i]) == o["flag"]
        ok = ok and res["scalings"][i].tobytes() == o["scalings"].tobytes()
        n_rec = len(o["records"])
        ok = ok and int(res["n_alignment"][i]) == n_rec
        if o["flag"] != 2:  # aligned: map, records, events per base
            ok = ok and res["events_per_base"][i] == o["events_per_base"]
            if records:
                off = int(res["pair_off"][i])
                got = res["records"][off:off + n_rec]
                ok = ok and [_rec_tuple(r) for r in got] == [_rec_tuple(r) for r in o["records"]]
            if base_map:
                ko = int(res["kmer_off"][i])
                nk = len(batch.read_seq(i)) - 5
                ok = ok and np.array_equal(res["base_to_event_map"][ko:ko + nk], o["map"])
        else:
            ok = ok and res["events_per_base"][i] == 0.0
        if not ok:
            bad.append((i, int(res["read_stat_flag"][i]), o["flag"], res["scalings"][i],
                        o["scalings"]))
    return bad


def test_postalign_golden_reads(ctx):
    b, gold, _ = golden_batch()
    res = run_post(ctx, b)
    assert compare_post(b, res) == []
    assert np.all(res["read_stat_flag"] == 0)


@pytest.mark.parametrize("form", [0, 1])
def test_postalign_mixed_qc_batch(ctx, form):
    b = qc_mixed_batch()
    res = run_post(ctx, b, form=form, aos=form == 1)
    assert compare_post(b, res) == []
    flags = set(res["read_stat_flag"].tolist())
    assert flags == {0, 1, 2, 4}, flags


def test_postalign_options(ctx):
    """scale_var off, and records / map not requested."""
    b = qc_mixed_batch(n=16, seed=4)
    res = run_post(ctx, b, scale_var=False, records=False, base_map=False)
    assert compare_post(b, res, scale_var=False, records=False, base_map=False) == []


def test_non_acgt_bases(ctx):
    """Bases other than A, C, G, T (N, lowercase, other bytes) rank as A in
    the reference; postalign, recalibration and MOM must agree with the
    oracle on such reads too."""
    import oracle
    from abea_amd import engine
    b = qc_mixed_batch(n=12, seed=31)
    rng = np.random.default_rng(31)
    seq = b.seq.copy()
    pos = rng.choice(seq.shape[0], seq.shape[0] // 50, replace=False)
    seq[pos] = rng.choice(np.frombuffer(b"NacgtX@\x00", np.uint8), pos.shape[0])
    b.seq = seq
    res = run_post(ctx, b)
    assert compare_post(b, res) == []
    db = engine.DeviceBatch.upload(b, 0)
    plan = engine.Plan(ctx, db, 1)
    got = plan.estimate_scalings()[:b.n_reads].cpu().numpy()
    plan.close()
    want = np.stack([oracle.estimate_scalings_mom(b.read_seq(i), b.read_events(i), b.model)
                     for i in range(b.n_reads)])
    assert got.tobytes() == want.tobytes()


def test_mom_golden_and_synthetic(ctx):
    """MOM on the GPU gives f5c's dumped scalings bit for bit, and the
    oracle's on synthetic reads."""
    import oracle
    from abea_amd import engine, synth
    b, _, _ = golden_batch()
    for batch, expect in ((b, b.scalings()), (synth.workload("cpu64x2kb", seed=9, n=16), None)):
        db = engine.DeviceBatch.upload(batch, 0)
        plan = engine.Plan(ctx, db, 1)
        got = plan.estimate_scalings()[:batch.n_reads].cpu().numpy()
        plan.close()
        if expect is None:
            expect = np.stack([oracle.estimate_scalings_mom(batch.read_seq(i),
                                                            batch.read_events(i), batch.model)
                               for i in range(batch.n_reads)])
        assert got.tobytes() == expect.tobytes()


def test_mom_mixed_lengths_and_serial_sums(ctx):
    """MOM on reads of 2 to 12 kb, some of whose event or level sums fail the
    exactness test (a 1e-30 event, a k-mer of level 1e-30) and so take the
    serial pass, among reads whose sums run in parallel.  The batched plan
    (queue longest first) and the host entry point (its own order) both
    match the oracle bit for bit."""
    import oracle
    from abea_amd import engine, synth
    L, lib = _lib()
    lengths = np.array([12000, 7500, 7000, 6800, 6600, 4000, 3800] + [2000 + 37 * i
                                                                      for i in range(23)])
    b = synth.generate(lengths, seed=41, skip=0.015)
    model = b.model.copy()
    model[0, 0] = np.float32(1e-30)  # AAAAAA
    b.model = model
    ev = b.ev_mean.copy()
    for i in (2, 9, 20):  # one tiny event
        ev[int(b.ev_ptr[i]) + 5] = np.float32(1e-30)
    b.ev_mean = ev
    seq = b.seq.copy()
    for i in (4, 13):  # AAAAAA
        p = int(b.seq_ptr[i]) + 100
        seq[p:p + 6] = np.frombuffer(b"AAAAAA", np.uint8)
    b.seq = seq
    n = b.n_reads
    want = np.stack([oracle.estimate_scalings_mom(b.read_seq(i), b.read_events(i), b.model)
                     for i in range(n)])
    plan = engine.Plan(ctx, engine.DeviceBatch.upload(b, 0), 1)
    got = plan.estimate_scalings()[:n].cpu().numpy()
    plan.close()
    assert got.tobytes() == want.tobytes()

    ev_all = b.events_aos()
    seqs = [C.create_string_buffer(b.read_seq(i), len(b.read_seq(i)) + 1) for i in range(n)]
    read = (C.c_char_p * n)(*[C.cast(s, C.c_char_p) for s in seqs])
    read_len = np.ascontiguousarray(b.read_len, np.int32)
    ets, keeps = [], []
    for i in range(n):
        et, k = _et(ev_all[int(b.ev_ptr[i]):int(b.ev_ptr[i]) + int(b.n_events[i])])
        ets.append(et)
        keeps.append(k)
    et_arr = (L.EventTable * n)(*ets)
    scal = (L.ScalingsT * n)()
    vp = lambda x: C.cast(x, C.c_void_p)  # noqa: E731
    mdl = np.ascontiguousarray(b.model, np.float32)
    assert lib.abea_estimate_scalings_db(n, vp(read), read_len.ctypes.data, vp(et_arr),
                                         mdl.ctypes.data, vp(scal)) == 0
    got2 = np.array([[s.scale, s.shift, s.var, s.log_var] for s in scal], np.float32)
    assert got2.tobytes() == want.tobytes()


def _first_diff(got: bytes, want: bytes, item: int = 32):
    """None when equal, else (index, got item, want item) of the first
    differing record."""
    if got == want:
        return None
    n = min(len(got), len(want))
    for i in range(0, n, item):
        if got[i:i + item] != want[i:i + item]:
            return (i // item, got[i:i + item].hex(), want[i:i + item].hex(), len(got), len(want))
    return (n // item, len(got), len(want))


def _lib():
    from abea_amd import _lib as L
    return L, L.lib()


def _et(ev_aos):
    L, _ = _lib()
    arr = (L.EventT * max(len(ev_aos), 1))()
    for j, e in enumerate(ev_aos):
        arr[j] = L.EventT(int(e["start"]), float(e["length"]), float(e["mean"]),
                          float(e["stdv"]))
    return L.EventTable(len(ev_aos), 0, len(ev_aos), C.cast(arr, C.POINTER(L.EventT))), arr


def test_per_read_dropins(ctx):
    """abea_estimate_scalings_using_mom / abea_postalign /
    abea_recalibrate_model over host buffers, against the oracle."""
    import oracle
    L, lib = _lib()
    b, gold, _ = golden_batch()
    model = np.ascontiguousarray(b.model, np.float32)
    ev_all = b.events_aos()
    for i in (0, 5, 17):
        seq = b.read_seq(i)
        ev = ev_all[int(b.ev_ptr[i]):int(b.ev_ptr[i]) + int(b.n_events[i])]
        et, keep = _et(ev)
        sc = lib.abea_estimate_scalings_using_mom(seq, len(seq), model.ctypes.data, et)
        want = oracle.estimate_scalings_mom(seq, ev["mean"], model)
        assert np.array([sc.scale, sc.shift, sc.var, sc.log_var], np.float32).tobytes() == \
            want.tobytes()

        pairs = np.ascontiguousarray(gold[i], np.int32)
        nk = len(seq) - 5
        rec = np.zeros(pairs.shape[0], np.dtype((np.void, 32)))
        bmap = np.zeros((nk, 2), np.int32)
        epb = C.c_double()
        n = lib.abea_postalign(rec.ctypes.data, bmap.ctypes.data, C.byref(epb), seq, nk,
                               pairs.ctypes.data, pairs.shape[0])
        w_rec, w_map, w_epb = oracle.postalign(seq, pairs)
        assert n == len(w_rec), L.lib().abea_last_error()
        assert _first_diff(rec[:n].tobytes(), w_rec.tobytes()) is None, \
            (i, _first_diff(rec[:n].tobytes(), w_rec.tobytes()))
        assert np.array_equal(bmap, w_map) and epb.value == w_epb

        s = L.ScalingsT(*[float(v) for v in b.scalings()[i]])
        ok = lib.abea_recalibrate_model(model.ctypes.data, et, C.byref(s), rec.ctypes.data, n,
                                        True)
        w_ok, w_sc = oracle.recalibrate(model, ev["mean"], b.scalings()[i], w_rec, True)
        assert ok == w_ok
        assert np.array([s.scale, s.shift, s.var, s.log_var], np.float32).tobytes() == \
            w_sc.tobytes()
        del keep


def _boundary_paths():
    """Paths whose length or whose runs of U and L steps straddle the walk's
    2048-pair work-item boundaries, over sequences with homopolymer stretches
    (repeated k-mers: 'E' records that start a new k-mer)."""
    rng = np.random.default_rng(77)
    out = []
    for n_target, pattern in ((2047, None), (2048, None), (2049, None), (4100, "L"),
                              (4100, "U"), (6200, "mix")):
        moves = list(rng.choice(3, n_target - 1, p=[0.5, 0.4, 0.1]))
        for b in (2048, 4096):
            if b - 3 < len(moves) and pattern in ("L", "U", "mix"):
                mv = {"L": 2, "U": 1}.get(pattern, 2)
                for j in range(b - 4, min(b + 3, len(moves))):
                    moves[j] = mv
                if pattern == "mix" and b + 5 < len(moves):
                    moves[b - 1], moves[b] = 1, 2
        k = e = 0
        pairs = [(0, 0)]
        for mv in moves:
            if mv == 0:
                k, e = k + 1, e + 1
            elif mv == 1:
                e += 1
            else:
                k += 1
            pairs.append((k, e))
        L = k + 6
        seq = bytearray(rng.choice(np.frombuffer(b"ACGT", np.uint8), L).tobytes())
        for start in range(0, L - 12, 300):  # homopolymer stretches
            seq[start:start + 9] = b"A" * 9
        out.append((bytes(seq), np.array(pairs, np.int32)))
    return out


def test_postalign_segment_boundaries(ctx):
    import oracle
    L, lib = _lib()
    for seq, pairs in _boundary_paths():
        nk = len(seq) - 5
        rec = np.zeros(pairs.shape[0], np.dtype((np.void, 32)))
        bmap = np.zeros((nk, 2), np.int32)
        epb = C.c_double()
        n = lib.abea_postalign(rec.ctypes.data, bmap.ctypes.data, C.byref(epb), seq, nk,
                               pairs.ctypes.data, pairs.shape[0])
        w_rec, w_map, w_epb = oracle.postalign(seq, pairs)
        assert n == len(w_rec), (pairs.shape[0], n, len(w_rec), lib.abea_last_error())
        assert _first_diff(rec[:n].tobytes(), w_rec.tobytes()) is None, pairs.shape[0]
        assert np.array_equal(bmap, w_map) and epb.value == w_epb


def test_per_read_dropin_errors(ctx):
    L, lib = _lib()
    seq = b"ACGTACGTACGTACGT"
    nk = len(seq) - 5
    rec = np.zeros(8, np.dtype((np.void, 32)))
    bmap = np.zeros((nk, 2), np.int32)
    epb = C.c_double()
    # not a path: the event index goes back
    bad = np.array([[0, 0], [1, 1], [2, 0]], np.int32)
    assert lib.abea_postalign(rec.ctypes.data, bmap.ctypes.data, C.byref(epb), seq, nk,
                              bad.ctypes.data, 3) == -1
    assert b"alignment path" in lib.abea_last_error()
    # k-mer out of range
    far = np.array([[0, 0], [1, 1], [nk, 2]], np.int32)
    assert lib.abea_postalign(rec.ctypes.data, bmap.ctypes.data, C.byref(epb), seq, nk,
                              far.ctypes.data, 3) == -1
    # no pairs
    assert lib.abea_postalign(rec.ctypes.data, bmap.ctypes.data, C.byref(epb), seq, nk,
                              far.ctypes.data, 0) == -1
    # an 'M' record whose event is outside the event table
    b, _, _ = golden_batch()
    model = np.ascontiguousarray(b.model, np.float32)
    ev = b.events_aos()[:10]
    et, keep = _et(ev)
    recs = np.zeros(300, np.dtype((np.void, 32))).view(np.uint8).reshape(300, 32)
    recs[:, 0:6] = np.frombuffer(b"ACGTAC", np.uint8)
    recs[:, 16:20] = np.array([50], np.int32).view(np.uint8)
    recs[:, 28] = ord("M")
    s = L.ScalingsT(1, 0, 1, 0)
    assert not lib.abea_recalibrate_model(model.ctypes.data, et, C.byref(s), recs.ctypes.data,
                                          300, True)
    assert b"event_idx" in lib.abea_last_error()
    del keep


def test_db_dropins(ctx):
    """abea_estimate_scalings_db and abea_postalign_db over f5c's db_t
    arrays, against the oracle's per-read flow."""
    import oracle
    L, lib = _lib()
    b = qc_mixed_batch(n=12, seed=8)
    n = b.n_reads
    model = np.ascontiguousarray(b.model, np.float32)
    ev_all = b.events_aos()
    seqs = [C.create_string_buffer(b.read_seq(i), len(b.read_seq(i)) + 1) for i in range(n)]
    read = (C.c_char_p * n)(*[C.cast(s, C.c_char_p) for s in seqs])
    read_len = np.ascontiguousarray(b.read_len, np.int32)
    ets, keeps = [], []
    for i in range(n):
        et, k = _et(ev_all[int(b.ev_ptr[i]):int(b.ev_ptr[i]) + int(b.n_events[i])])
        ets.append(et)
        keeps.append(k)
    et_arr = (L.EventTable * n)(*ets)
    scal = (L.ScalingsT * n)()
    vp = lambda x: C.cast(x, C.c_void_p)  # noqa: E731
    assert lib.abea_estimate_scalings_db(n, vp(read), read_len.ctypes.data, vp(et_arr),
                                         model.ctypes.data, vp(scal)) == 0
    for i in range(n):
        want = oracle.estimate_scalings_mom(b.read_seq(i), b.read_events(i), model)
        got = np.array([scal[i].scale, scal[i].shift, scal[i].var, scal[i].log_var], np.float32)
        assert got.tobytes() == want.tobytes()

    # align with the generating scalings, then the post-align flow
    scal = (L.ScalingsT * n)(*[L.ScalingsT(*[float(v) for v in row]) for row in b.scalings()])
    pairs = (C.POINTER(L.AlignedPair) * n)()
    npairs = np.zeros(n, np.int32)
    assert lib.abea_align_db(n, vp(read), read_len.ctypes.data, vp(et_arr), model.ctypes.data,
                             vp(scal), vp(pairs), npairs.ctypes.data) == 0
    ea = (C.c_void_p * n)()
    n_ea = np.zeros(n, np.int32)
    epb = np.zeros(n, np.float64)
    bmaps = (C.c_void_p * n)()
    flags = np.zeros(n, np.int32)
    assert lib.abea_postalign_db(n, vp(read), read_len.ctypes.data, vp(et_arr),
                                 model.ctypes.data, vp(pairs), npairs.ctypes.data, vp(scal),
                                 vp(ea), n_ea.ctypes.data, epb.ctypes.data, vp(bmaps),
                                 flags.ctypes.data) == 0, lib.abea_last_error()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]  # pointers, not C ints
    libc.free.restype = None
    for i in range(n):
        o = oracle.process_read(b.read_seq(i), b.read_events(i), model, b.scalings()[i])
        assert flags[i] == o["flag"], i
        got_sc = np.array([scal[i].scale, scal[i].shift, scal[i].var, scal[i].log_var],
                          np.float32)
        assert got_sc.tobytes() == o["scalings"].tobytes(), i
        assert n_ea[i] == len(o["records"])
        if npairs[i] > 0:
            got = np.ctypeslib.as_array(C.cast(ea[i], C.POINTER(C.c_uint8)),
                                        (int(n_ea[i]) * 32,))
            assert _first_diff(got.tobytes(), o["records"].tobytes()) is None, \
                (i, _first_diff(got.tobytes(), o["records"].tobytes()))
            nk = len(b.read_seq(i)) - 5
            gm = np.ctypeslib.as_array(C.cast(bmaps[i], C.POINTER(C.c_int32)), (nk * 2,))
            assert np.array_equal(gm.reshape(nk, 2), o["map"])
            assert epb[i] == o["events_per_base"]
            libc.free(ea[i])
            libc.free(bmaps[i])
        else:
            assert not ea[i] and not bmaps[i] and epb[i] == 0.0
        libc.free(C.cast(pairs[i], C.c_void_p))
