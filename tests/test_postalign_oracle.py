"""Pinning the oracle's restatement of f5c's steps either side of align():
estimate_scalings_using_mom (against the scalings f5c itself dumped),
postalign (against a hand-worked path and an event-assignment definition
independent of f5c's loop) and recalibrate_model (against a least-squares
solve by numpy and its threshold).  Each check also runs on a planted
mistake in the oracle to show that it would see it."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REFERENCE_DUMP, ROOT

import oracle


def golden():
    from abea_amd import load_golden_subset
    return load_golden_subset(GOLDEN)


def _mutant(name):
    so = os.path.join(ROOT, "oracle", "_build", f"liborc_{name.lower()}.so")
    flags = ["-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", f"-D{name}"]
    subprocess.run(["gcc"] + flags + ["-o", so, os.path.join(ROOT, "oracle", "abea_oracle.c"),
                                      os.path.join(ROOT, "oracle", "postalign.c"),
                                      os.path.join(ROOT, "oracle", "getevents.c"), "-lm"],
                   check=True)
    return so


def _with_lib(so, fn):
    saved_lib, saved_path = oracle._lib, oracle.LIB_PATH
    oracle._lib, oracle.LIB_PATH = None, so
    try:
        return fn()
    finally:
        oracle._lib, oracle.LIB_PATH = saved_lib, saved_path


# ---------------------------------------------------------------- MOM

def _mom_mismatches(b):
    bad = 0
    sc = b.scalings()
    for i in range(b.n_reads):
        got = oracle.estimate_scalings_mom(b.read_seq(i), b.read_events(i), b.model)
        if not (got[0] == sc[i, 0] and got[1] == sc[i, 1] and got[2] == 0 and got[3] == 0):
            bad += 1
    return bad


def test_mom_equals_f5c_dumped_scalings():
    """f5c passes estimate_scalings_using_mom()'s result to align(); the dump
    holds it (var and log_var 0) for every read."""
    b, _, _ = golden()
    assert np.all(b.var == 0) and np.all(b.log_var == 0)
    assert _mom_mismatches(b) == 0


def test_mom_on_all_143_dump_reads():
    if not os.path.isdir(REFERENCE_DUMP):
        pytest.skip("reference dump not mounted")
    from abea_amd import dump
    db = dump.load_batch(os.path.join(REFERENCE_DUMP, "0"), with_outputs=False)
    assert _mom_mismatches(dump.to_engine_batch(db)) == 0


def test_mom_catches_planted_mistake():
    b, _, _ = golden()
    assert _with_lib(_mutant("ORC_MUTATE_MOM_FLOAT_RESIDUAL"), lambda: _mom_mismatches(b)) > 0


# ---------------------------------------------------------------- postalign

# A hand-worked path over 12 bases (7 k-mers): D, U, L, L, D, U, D, D.
HAND_SEQ = b"ACGTAAAAAAAC"
HAND_PAIRS = [(0, 0), (1, 1), (1, 2), (2, 2), (3, 2), (4, 3), (4, 4), (5, 5), (6, 6)]
# event e belongs to the k-mer of its first pair: 0->0, 1->1, 2->1, 3->4,
# 4->4, 5->5, 6->6; k-mers 2 and 3 are only passed by k-mer skips.
HAND_MAP = [(0, 0), (1, 2), (-1, -1), (-1, -1), (3, 4), (5, 5), (6, 6)]
# k-mers 4 and 5 are both AAAAAA: the record of k-mer 5 repeats the rank -> E
HAND_RECORDS = [(0, 0, b"M"), (1, 1, b"M"), (1, 2, b"E"), (4, 3, b"M"), (4, 4, b"E"),
                (5, 5, b"E"), (6, 6, b"M")]


def _hand_ok():
    rec, bmap, epb = oracle.postalign(HAND_SEQ, np.array(HAND_PAIRS, np.int32))
    return (epb == 6 / 7 and [tuple(x) for x in bmap] == HAND_MAP
            and [(int(r["ref_position"]), int(r["event_idx"]), r["hmm_state"]) for r in rec]
            == HAND_RECORDS
            and all(r["ref_kmer"] == HAND_SEQ[r["ref_position"]:r["ref_position"] + 6]
                    and r["model_kmer"] == r["ref_kmer"] and r["read_idx"] == -1
                    and r["rc"] == 0 for r in rec))


def test_postalign_hand_worked_path():
    assert _hand_ok()


def expected_postalign(seq: bytes, pairs):
    """Definition by event assignment: every event on the path belongs to the
    k-mer of its first pair; records are the events in order."""
    n_k = len(seq) - 5
    owner = {}
    for k, e in pairs:
        owner.setdefault(int(e), int(k))
    events = sorted(owner)
    bmap = np.full((n_k, 2), -1, np.int32)
    for e in events:
        k = owner[e]
        if bmap[k, 0] < 0:
            bmap[k, 0] = e
        bmap[k, 1] = max(bmap[k, 1], e)
    rec, prev = [], None
    for e in events:
        r = oracle.kmer_rank(seq[owner[e]:owner[e] + 6])
        rec.append((owner[e], e, b"M" if r != prev else b"E"))
        prev = r
    es = [int(e) for _, e in pairs]
    return rec, bmap, (max(es) - min(es)) / n_k


def _check_against_definition(seq, pairs):
    rec, bmap, epb = oracle.postalign(seq, pairs)
    xr, xb, xe = expected_postalign(seq, pairs)
    return (np.array_equal(bmap, xb) and epb == xe
            and [(int(r["ref_position"]), int(r["event_idx"]), r["hmm_state"]) for r in rec] == xr)


def _path_cases():
    b, gold, _ = golden()
    cases = [(b.read_seq(i), gold[i]) for i in range(b.n_reads)]
    # random paths with frequent skips and stays over a low-complexity
    # sequence (repeated k-mers exercise the E state on a new k-mer)
    rng = np.random.default_rng(5)
    for t in range(30):
        L = int(rng.integers(7, 400))
        alphabet = np.frombuffer(b"AAAC" if t % 3 == 0 else b"ACGT", np.uint8)
        seq = rng.choice(alphabet, L).tobytes()
        n_k = L - 5
        k, e = int(rng.integers(0, 3)) % n_k, int(rng.integers(0, 3))
        pairs = [(k, e)]
        while k < n_k - 1:
            mv = rng.choice(3, p=[0.5, 0.3, 0.2])  # D, U, L
            if mv == 0:
                k, e = k + 1, e + 1
            elif mv == 1:
                e += 1
            else:
                k += 1
            pairs.append((k, e))
        cases.append((seq, np.array(pairs, np.int32)))
    return cases


def test_postalign_matches_event_assignment_definition():
    bad = [i for i, (s, p) in enumerate(_path_cases()) if not _check_against_definition(s, p)]
    assert bad == []


@pytest.mark.parametrize("mutant", ["ORC_MUTATE_POSTALIGN_EVERY_PAIR",
                                    "ORC_MUTATE_POSTALIGN_STATE_BY_POSITION"])
def test_postalign_checks_catch_planted_mistakes(mutant):
    def count():
        n = sum(not _check_against_definition(s, p) for s, p in _path_cases())
        return n + (0 if _hand_ok() else 1)
    assert _with_lib(_mutant(mutant), count) > 0


# ---------------------------------------------------------------- recalibrate

def _records(n_m, n_e_extra, seed):
    """n_m 'M' records on distinct random k-mers with event means from a
    known (scale, shift) plus noise, interleaved with n_e_extra 'E' records
    whose events are far off (they must not matter)."""
    rng = np.random.default_rng(seed)
    model = np.zeros((4096, 3), np.float32)
    model[:, 0] = rng.uniform(60, 130, 4096)
    model[:, 1] = rng.uniform(1.5, 3.5, 4096)
    model[:, 2] = np.log(model[:, 1])
    bases = np.frombuffer(b"ACGT", np.uint8)
    kmers = [rng.choice(bases, 6).tobytes() for _ in range(n_m)]
    ranks = np.array([oracle.kmer_rank(k) for k in kmers])
    scale, shift = 1.07, -3.5
    mu, sd = model[ranks, 0].astype(np.float64), model[ranks, 1].astype(np.float64)
    x = (scale * mu + shift + rng.normal(0, 1.3, n_m) * sd).astype(np.float32)
    ev = list(x)
    rec = []
    for i in range(n_m):
        rec.append((kmers[i], i, b"M"))
    for j in range(n_e_extra):
        ev.append(np.float32(1000.0 + j))
        rec.insert(int(rng.integers(0, len(rec) + 1)), (kmers[j % n_m], n_m + j, b"E"))
    arr = np.zeros(len(rec), oracle.EVENT_ALIGNMENT_DTYPE)
    for i, (k, e, st) in enumerate(rec):
        arr[i]["ref_kmer"] = k
        arr[i]["model_kmer"] = k
        arr[i]["ref_position"] = i
        arr[i]["read_idx"] = -1
        arr[i]["event_idx"] = e
        arr[i]["hmm_state"] = st
    return model, np.array(ev, np.float32), arr, (mu, sd, x.astype(np.float64))


def _lstsq(mu, sd, x):
    w = 1.0 / sd
    A = np.stack([w, mu * w], 1)
    (shift, scale), *_ = np.linalg.lstsq(A, x * w, rcond=None)
    var = np.sqrt(np.mean(((x - shift - scale * mu) / sd) ** 2))
    return shift, scale, var


def _ulps(a, b):
    a, b = np.float32(a), np.float32(b)
    return abs(int(a.view(np.int32)) - int(b.view(np.int32)))


def _recal_ok(seed):
    model, ev, rec, (mu, sd, x) = _records(300, 120, seed)
    ok, sc = oracle.recalibrate(model, ev, np.array([1, 0, 1, 0], np.float32), rec, True)
    shift, scale, var = _lstsq(mu, sd, x)
    return (ok and _ulps(sc[1], shift) <= 2 and _ulps(sc[0], scale) <= 2
            and _ulps(sc[2], var) <= 2 and _ulps(sc[3], np.log(var)) <= 2)


def test_recalibrate_is_weighted_least_squares():
    assert all(_recal_ok(seed) for seed in range(5))


def test_recalibrate_catches_planted_mistake():
    assert _with_lib(_mutant("ORC_MUTATE_RECAL_ALL_RECORDS"),
                     lambda: sum(not _recal_ok(s) for s in range(5))) > 0


def test_recalibrate_threshold_and_scale_var():
    start = np.array([1.5, 2.5, 3.5, 4.5], np.float32)
    for n_m, want in ((199, False), (200, True)):
        model, ev, rec, _ = _records(n_m, 50, 11)
        ok, sc = oracle.recalibrate(model, ev, start, rec, True)
        assert ok == want
        if not want:
            assert np.array_equal(sc, start)
    model, ev, rec, (mu, sd, x) = _records(250, 0, 12)
    ok, sc = oracle.recalibrate(model, ev, start, rec, False)
    shift, scale, _ = _lstsq(mu, sd, x)
    assert ok and sc[2] == 1.0 and sc[3] == 0.0
    assert _ulps(sc[1], shift) <= 2 and _ulps(sc[0], scale) <= 2


def test_process_read_flow_on_golden_reads():
    """f5c's flow after align() on the real reads: every read aligns, so every
    read is recalibrated from its 'M' records and passes (var <= 2.5)."""
    b, gold, _ = golden()
    for i in range(b.n_reads):
        r = oracle.process_read(b.read_seq(i), b.read_events(i), b.model, b.scalings()[i])
        assert np.array_equal(r["align"]["pairs"], gold[i])
        assert r["calibrated"] and r["flag"] == 0
        assert 0.5 < r["scalings"][2] < 2.5
        assert len(r["records"]) == gold[i][-1, 1] - gold[i][0, 1] + 1
