"""Pins the CPU oracle (the checker of the HIP path) to the reference.

1. Golden vectors: f5c's own event_align_pairs for real reads
   (FPGA/code/bins/dump_small), bit-exact.
2. Planted mistakes: mutants of the oracle compiled on the CPU must fail the
   golden or brute-force checks (so those checks can see such mistakes).
3. Brute force: for n_events + n_kmers <= 49 the W=100 band covers the whole
   DP matrix, so ABEA equals a full-matrix Viterbi with the same recurrences;
   an independent full-matrix implementation below must agree exactly.
4. Closed forms and known answers: k-mer ranks, transition constants,
   emission at a = 0 and a = 1.
5. FP64 gradient against central finite differences with the band held fixed.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REFERENCE_DUMP, ROOT

import oracle

oracle.build()


def golden():
    from abea_amd import load_golden_subset
    return load_golden_subset(GOLDEN)


def run_oracle_on(b, i, **kw):
    return oracle.align(b.read_seq(i), b.read_events(i), b.model, float(b.scale[i]),
                        float(b.shift[i]), **kw)


def test_golden_subset_bit_exact():
    b, gold, _ = golden()
    for i in range(b.n_reads):
        r = run_oracle_on(b, i)
        assert r["n_pairs"] == gold[i].shape[0], i
        assert np.array_equal(r["pairs"], gold[i]), i


@pytest.mark.skipif(not os.path.isdir(REFERENCE_DUMP), reason="reference dump not mounted")
def test_full_golden_dump_143_reads():
    from abea_amd import dump
    bt = dump.load_batch(os.path.join(REFERENCE_DUMP, "0"))
    assert len(bt.reads) == 143
    for i, r in enumerate(bt.reads):
        res = oracle.align(r.seq, r.events["mean"], bt.model, float(r.scalings["scale"]),
                           float(r.scalings["shift"]))
        assert np.array_equal(res["pairs"], r.pairs), i


# ---------------------------------------------------------------- brute force
def full_matrix_viterbi(seq: bytes, ev: np.ndarray, model: np.ndarray, scale, shift):
    """Unbanded ABEA recurrences over the whole (n_e+1) x (n_k+1) matrix, written
    independently of the banded oracle: row -1/col -1 boundary (start 0, trim
    column lp_trim*(e+1)), f5c float cells with double transition sums, the
    reference's D<U<L tie order, first maximum in the end search."""
    n_k, n_e = len(seq) - 5, len(ev)
    lp_skip, lp_stay, lp_step, lp_trim = oracle.transition_constants(n_e, n_k)
    f32 = np.float32
    NEG = -math.inf
    S = [[NEG] * (n_k + 1) for _ in range(n_e + 1)]  # S[e+1][k+1]
    T = [[0] * (n_k + 1) for _ in range(n_e + 1)]
    S[0][0] = 0.0
    for e in range(n_e):
        S[e + 1][0] = float(f32(lp_trim * (e + 1)))
    for e in range(n_e):
        for k in range(n_k):
            rank = oracle.kmer_rank(seq[k:k + 6])
            em = float(oracle.emission(ev[e], model[rank], scale, shift))
            sd = float(f32(S[e][k] + lp_step + em))
            su = float(f32(S[e][k + 1] + lp_stay + em))
            sl = float(f32(S[e + 1][k] + lp_skip))
            best, fr = sd, 0
            if su >= best:
                best, fr = su, 1
            if sl >= best:
                best, fr = sl, 2
            S[e + 1][k + 1] = best
            T[e + 1][k + 1] = fr
    best, be = NEG, 0
    for e in range(n_e):
        s = float(f32(S[e + 1][n_k] + (n_e - e) * lp_trim))
        if s > best:
            best, be = s, e
    path, k, e = [], n_k - 1, be
    while k >= 0 and e >= 0:
        path.append((k, e))
        f = T[e + 1][k + 1]
        if f == 0:
            k, e = k - 1, e - 1
        elif f == 1:
            e -= 1
        else:
            k -= 1
    return best, np.array(path[::-1], np.int32).reshape(-1, 2)


def tiny_cases(n=40, seed=7, degenerate=False):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        n_k = int(rng.integers(1, 24))
        n_e = int(rng.integers(1, 49 - n_k + 1))
        seq = rng.choice(np.frombuffer(b"ACGT", np.uint8), n_k + 5).tobytes()
        if degenerate:
            # identical emissions everywhere: exact ties exercise the tie order
            ev = np.full(n_e, 100.0, np.float32)
        else:
            ev = rng.uniform(55, 135, n_e).astype(np.float32)
        out.append((seq, ev))
    return out


def flat_model():
    m = np.zeros((4096, 3), np.float32)
    m[:, 0] = 100.0
    m[:, 1] = 2.0
    m[:, 2] = np.float32(np.log(np.float32(2.0)))
    return m


@pytest.mark.parametrize("degenerate", [False, True])
def test_brute_force_full_matrix(degenerate):
    b, _, _ = golden()
    model = flat_model() if degenerate else b.model
    for seq, ev in tiny_cases(degenerate=degenerate, seed=3 if degenerate else 7):
        r = oracle.align(seq, ev, model, 1.0, 0.0)
        assert r["fills"] == len(ev) * (len(seq) - 5)  # the band covered every cell
        best, path = full_matrix_viterbi(seq, ev, model, 1.0, 0.0)
        assert np.float32(r["score"]) == np.float32(best)
        assert np.array_equal(r["path"], path)


# ------------------------------------------------------------ planted mistakes
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


@pytest.mark.parametrize("mutant", ["ORC_MUTATE_RANK_LSB_FIRST", "ORC_MUTATE_TIE_ORDER",
                                    "ORC_MUTATE_FLOAT_TRANSITIONS"])
def test_golden_catches_planted_mistake(mutant):
    b, gold, _ = golden()

    def count():
        return sum(not np.array_equal(run_oracle_on(b, i)["pairs"], gold[i])
                   for i in range(b.n_reads))
    assert _with_lib(_mutant(mutant), count) > 0


def end_tie_cases():
    """Tiny reads whose end cells tie: events sit on the model mean (a = 0) and
    the cached log-stdv is chosen so that one more 'stay' costs exactly one
    trim step (em + lp_stay == lp_trim in float), so several e give the same
    end score and only the first-maximum rule picks the reference's."""
    rng = np.random.default_rng(17)
    out = []
    for n_k, n_e in [(4, 20), (6, 30), (8, 25), (10, 38), (5, 40)]:
        lp_skip, lp_stay, lp_step, lp_trim = oracle.transition_constants(n_e, n_k)
        m = flat_model()
        c = np.float32(lp_trim - lp_stay)
        m[:, 2] = np.float32(np.float32(-0.918938) - c)
        seq = rng.choice(np.frombuffer(b"ACGT", np.uint8), n_k + 5).tobytes()
        out.append((seq, np.full(n_e, 100.0, np.float32), m))
    return out


@pytest.mark.parametrize("mutant", ["ORC_MUTATE_TIE_ORDER", "ORC_MUTATE_END_LAST_MAX"])
def test_brute_force_catches_planted_mistake(mutant):
    """Tie-order mistakes are invisible on smooth data; degenerate tiny cases
    with exact ties expose them against the full-matrix DP."""
    cases = [(s, e, flat_model()) for s, e in tiny_cases(degenerate=True, seed=3)]
    cases += end_tie_cases()
    expected = [full_matrix_viterbi(s, e, m, 1.0, 0.0)[1] for s, e, m in cases]
    # the unmutated oracle agrees with the brute force on all of them
    assert all(np.array_equal(oracle.align(s, e, m, 1.0, 0.0)["path"], p)
               for (s, e, m), p in zip(cases, expected))

    def count():
        return sum(not np.array_equal(oracle.align(s, e, m, 1.0, 0.0)["path"], p)
                   for (s, e, m), p in zip(cases, expected))
    assert _with_lib(_mutant(mutant), count) > 0


# -------------------------------------------------------- closed forms / KATs
def test_kmer_rank_known_answers():
    assert oracle.kmer_rank(b"AAAAAA") == 0
    assert oracle.kmer_rank(b"TTTTTT") == 4095
    assert oracle.kmer_rank(b"ACGTAC") == 0b000110110001  # 433: first base most significant
    assert oracle.kmer_rank(b"CAAAAA") == 1024
    assert oracle.kmer_rank(b"AAAAAC") == 1
    assert oracle.kmer_rank(b"NNNNNN") == 0  # non-ACGT counts as A
    from abea_amd import synth
    codes = np.array([0, 1, 2, 3, 0, 1, 3], np.uint8)
    assert list(synth.kmer_ranks(codes)) == [oracle.kmer_rank(b"ACGTAC"),
                                             oracle.kmer_rank(b"CGTACT")]


def test_transition_constants_closed_form():
    for n_e, n_k in [(100, 60), (3000, 1995), (7, 9)]:
        lp_skip, lp_stay, lp_step, lp_trim = oracle.transition_constants(n_e, n_k)
        epk = n_e / n_k
        assert lp_skip == math.log(1e-10)
        assert lp_trim == math.log(0.01)
        assert math.isclose(lp_stay, math.log(epk / (epk + 1)), rel_tol=1e-14)
        # the three transition probabilities sum to one
        assert math.isclose(math.exp(lp_skip) + math.exp(lp_stay) + math.exp(lp_step), 1.0,
                            rel_tol=1e-14)


def test_emission_special_points():
    row = np.array([80.0, 2.5, np.log(np.float32(2.5))], np.float32)
    c = np.float32(np.float32(-0.918938) - row[2])
    # a = 0: em is exactly the normalising constant minus log(sigma)
    assert oracle.emission(np.float32(80.0), row, 1.0, 0.0) == c
    # a = 1 (x = mean + sigma): em = c - 1/2 exactly in float
    assert oracle.emission(np.float32(82.5), row, 1.0, 0.0) == np.float32(c - np.float32(0.5))
    # scale and shift enter as scale*mean + shift
    assert oracle.emission(np.float32(90.0), row, 1.0, 10.0) == c
    assert oracle.emission(np.float32(160.0), row, 2.0, 0.0) == c


def test_fp64_oracle_close_to_f5c():
    """The FP64 oracle (error reference) and f5c's float semantics agree to the
    float rounding of the scores; the paths are the same on these reads."""
    b, _, _ = golden()
    for i in range(4):
        r32 = run_oracle_on(b, i)
        r64 = run_oracle_on(b, i, mode=oracle.MODE_FP64)
        assert abs(r32["score"] - r64["score"]) <= 1e-5 * abs(r64["score"])
        assert r32["path_len"] == r64["path_len"]


# ---------------------------------------------------------------- gradients
def _score64(seq, ev, model, scale, shift, moves):
    r = oracle.align(seq, ev, model, scale, shift, mode=oracle.MODE_FP64, replay_moves=moves)
    return r["score"], r["path"]


def _grad_case(b, i):
    seq, ev = b.read_seq(i), b.read_events(i)
    sc, sh = float(b.scale[i]), float(b.shift[i])
    base = oracle.align(seq, ev, b.model, sc, sh, mode=oracle.MODE_FP64)
    g = oracle.path_gradient(base["path"], base["from"], seq, ev, b.model, sc, sh)
    ranks = [oracle.kmer_rank(seq[k:k + 6]) for k, _ in base["path"]]
    top = sorted(set(ranks), key=ranks.count)[-3:]
    return seq, ev, sc, sh, base, g, top


def _perturbed(b, sc, sh, what, h, sign):
    """(model, scale, shift) with one parameter moved by sign*h (rounded to
    float32, as the inputs are float32), and the actual step taken."""
    model = b.model.copy()
    if what[0] == "shift":
        v = np.float32(sh + sign * h)
        return model, sc, float(v), float(v) - sh
    if what[0] == "scale":
        v = np.float32(sc + sign * h)
        return model, float(v), sh, float(v) - sc
    r, col = what[1], what[2]
    v = np.float32(model[r, col] + sign * h)
    step = float(v) - float(model[r, col])
    model[r, col] = v
    if col == 1:  # level_log_stdv = log(level_stdv)
        model[r, 2] = np.float32(np.log(np.float64(v)))
    return model, sc, sh, step


def _params(top):
    return [("shift",), ("scale",)] + [("mu", r, 0) for r in top] + [("sigma", r, 1) for r in top]


def _analytic(g, what):
    g_read, g_mu, g_sg = g
    if what[0] == "shift":
        return g_read[0]
    if what[0] == "scale":
        return g_read[1]
    return (g_mu if what[2] == 0 else g_sg)[what[1]]


def test_gradient_vs_fd_on_the_held_path():
    """Central differences (h = 1e-3) of the FP64 score of the traced path
    against the analytic gradient: pins every emission derivative."""
    from abea_amd import synth
    b = synth.workload("cpu64x2kb", seed=9, n=2)
    seq, ev, sc, sh, base, g, top = _grad_case(b, 0)
    assert g[0][2] == 0.0  # var never enters the emission (GPU/align_cuda.cl:95)
    for what in _params(top):
        mp, scp, shp, stp = _perturbed(b, sc, sh, what, 1e-3, +1)
        mm, scm, shm, stm = _perturbed(b, sc, sh, what, 1e-3, -1)
        fp = oracle.path_score(base["path"], base["from"], seq, ev, mp, scp, shp)
        fm = oracle.path_score(base["path"], base["from"], seq, ev, mm, scm, shm)
        fd = (fp - fm) / (stp - stm)
        an = _analytic(g, what)
        tol = 1e-3 * abs(an) + 1e-6
        if what[0] == "sigma":
            # the stored log-stdv is float32: each perturbed cell carries up to
            # half an ulp of log(sigma) of error, divided by the step
            n_cells = sum(1 for (k, _), f in zip(base["path"], base["from"])
                          if f != 2 and oracle.kmer_rank(seq[k:k + 6]) == what[1])
            lam = abs(float(b.model[what[1], 2]))
            tol += n_cells * lam * 2.0 ** -24 * 2 / (stp - stm)
        assert abs(fd - an) <= tol, (what, fd, an, tol)


def test_path_score_equals_dp_score():
    b, _, _ = golden()
    for i in range(3):
        r = run_oracle_on(b, i, mode=oracle.MODE_FP64)
        s = oracle.path_score(r["path"], r["from"], b.read_seq(i), b.read_events(i), b.model,
                              b.scale[i], b.shift[i])
        assert abs(s - r["score"]) <= 1e-9 * abs(r["score"])


@pytest.mark.parametrize("h", [1e-3, 1e-5])
def test_gradient_vs_fd_band_fixed(h):
    """Central differences of the FP64 DP score with the band held fixed
    (replayed moves).  The best path is piecewise constant in the parameters:
    wherever it does not switch inside [-h, h] the FD must match to 1e-6; at
    h = 1e-5 it must not switch for any tested parameter."""
    from abea_amd import synth
    b = synth.workload("cpu64x2kb", seed=9, n=2)
    seq, ev, sc, sh, base, g, top = _grad_case(b, 0)
    smooth = 0
    params = _params(top)
    for what in params:
        mp, scp, shp, stp = _perturbed(b, sc, sh, what, h, +1)
        mm, scm, shm, stm = _perturbed(b, sc, sh, what, h, -1)
        fp, pp = _score64(seq, ev, mp, scp, shp, base["moves"])
        fm, pm = _score64(seq, ev, mm, scm, shm, base["moves"])
        if not (np.array_equal(pp, base["path"]) and np.array_equal(pm, base["path"])):
            continue
        smooth += 1
        fd = (fp - fm) / (stp - stm)
        an = _analytic(g, what)
        tol = 1e-6 * abs(an) + 1e-6
        if what[0] == "sigma":
            n_cells = sum(1 for (k, _), f in zip(base["path"], base["from"])
                          if f != 2 and oracle.kmer_rank(seq[k:k + 6]) == what[1])
            tol += n_cells * abs(float(b.model[what[1], 2])) * 2.0 ** -24 * 2 / (stp - stm)
        assert abs(fd - an) <= tol, (what, fd, an, tol)
    if h <= 1e-5:
        assert smooth == len(params)
    else:
        assert smooth >= 1


# --------------------------------------------------------------------- QC
def _emission_f32(x, mu, sd, lsd, scale, shift):
    """f5c's emission in float32 steps (GPU/align_cuda.cl:511-528)."""
    f = np.float32
    m = f(f(scale) * f(mu)) + f(shift)
    a = (f(x) - m) / f(sd)
    return (f(-0.918938) - f(lsd)) + (f(-0.5) * a) * a


def _rank(s):
    code = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}
    r = 0
    for c in s:
        r = (r << 2) | code.get(c, 0)
    return r


def _qc_stats(seq, ev, model, scale, shift, path, frm):
    """QC statistics of a traced path, restated from the reference
    (GPU/align_cuda.cl:755-882): double sum of float emissions over every
    traced cell in traceback order, longest run of L moves, span."""
    total = 0.0
    for k, e in path[::-1]:
        mu, sd, lsd = model[_rank(seq[k:k + 6])]
        total += float(_emission_f32(ev[e], mu, sd, lsd, scale, shift))
    run = gap = 0
    for f in frm[::-1]:
        run = run + 1 if f == 2 else 0
        gap = max(gap, run)
    n_k = len(seq) - 5
    spanned = path[0][0] == 0 and path[-1][0] == n_k - 1
    return total / len(path), gap, spanned


def test_qc_rules_against_path_statistics():
    """The oracle's QC decision (avg log-emission >= -5, spanned, max gap <=
    50; GPU/align_cuda.cl:134,139,873) equals the rule applied to statistics
    recomputed from its path, and the average and gap rules both pass and
    fail somewhere in the set.  (A traced path is always spanned: it starts
    at the last k-mer, and the only finite predecessors of a cell on event 0
    lead to k-mer 0 or to the trim column, so the span test only guards
    against malformed paths.)"""
    b, _, _ = golden()
    rng = np.random.default_rng(23)
    cases = [(b.read_seq(i), b.read_events(i), b.model, float(b.scale[i]), float(b.shift[i]))
             for i in range(4)]
    from abea_amd import synth
    s = synth.workload("cpu64x2kb", seed=5, n=4)
    cases += [(s.read_seq(i), s.read_events(i), s.model, float(s.scale[i]), float(s.shift[i]))
              for i in range(4)]
    # few events for many k-mers: long runs of L moves (gaps > 50)
    for n_k, n_e in [(40, 6), (300, 5), (60, 12), (250, 3)]:
        sq = rng.choice(np.frombuffer(b"ACGT", np.uint8), n_k + 5).tobytes()
        cases.append((sq, rng.uniform(60, 130, n_e).astype(np.float32), b.model, 1.0, 0.0))
    fired = {"avg": set(), "gap": set(), "span": set()}
    for seq, ev, model, scale, shift in cases:
        r = oracle.align(seq, ev, model, scale, shift)
        path, frm = r["path"], r["from"]
        assert len(path) == r["path_len"] > 0
        avg, gap, spanned = _qc_stats(seq, ev, model, scale, shift, path, frm)
        assert avg == r["avg_log_emission"]
        assert gap == r["max_gap"] and bool(spanned) == bool(r["spanned"])
        ok = avg >= -5.0 and spanned and gap <= 50
        assert bool(r["qc_pass"]) == ok
        assert r["n_pairs"] == (len(path) if ok else 0)
        fired["avg"].add(avg < -5.0)
        fired["gap"].add(gap > 50)
        fired["span"].add(not spanned)
    assert fired["avg"] == {True, False} and fired["gap"] == {True, False}, fired
    assert fired["span"] == {False}


# ------------------------------------------------------ batch driver, fills
def test_batch_driver_equals_per_read():
    """The OpenMP batch driver (CPU baseline and bulk parity reference) gives
    each read exactly what the per-read oracle gives, in both modes."""
    from abea_amd import synth
    b = synth.workload("cpu64x2kb", seed=17, n=6)
    for mode in (oracle.MODE_F5C, oracle.MODE_FP64):
        ob = oracle.align_batch(b, n_threads=3, mode=mode, want_path=True, want_pairs=True)
        for i in range(b.n_reads):
            r = run_oracle_on(b, i, mode=mode)
            off, n = int(ob["pair_ptr"][i]), int(ob["path_len"][i])
            assert n == r["path_len"] and ob["n_pairs"][i] == r["n_pairs"]
            assert np.array_equal(ob["path"][off:off + n], r["path"])
            assert np.array_equal(ob["from"][off:off + n], r["from"])
            assert np.array_equal(ob["pairs"][off:off + r["n_pairs"]], r["pairs"])
            assert ob["score"][i] == r["score"] and ob["avg"][i] == r["avg_log_emission"]
            assert ob["fills"][i] == r["fills"]


def test_fills_from_moves():
    """Cells computed per read, recomputed from packed band moves, equal the
    fill's own count; when the band covers the whole matrix (n_e + n_k <=
    49) every matrix cell is filled exactly once: fills = n_e * n_k."""
    from abea_amd import engine, synth
    b = synth.workload("cpu64x2kb", seed=19, n=5)
    rng = np.random.default_rng(19)
    seqs = [b.read_seq(i) for i in range(b.n_reads)]
    evs = [b.read_events(i) for i in range(b.n_reads)]
    for n_k, n_e in [(20, 25), (1, 40), (30, 3), (44, 5)]:
        seqs.append(rng.choice(np.frombuffer(b"ACGT", np.uint8), n_k + 5).tobytes())
        evs.append(rng.uniform(60, 130, n_e).astype(np.float32))
    res = [oracle.align(s, e, b.model, 1.0, 0.0) for s, e in zip(seqs, evs)]
    rl = np.array([len(s) for s in seqs], np.int32)
    ne = np.array([len(e) for e in evs], np.int32)
    words = engine.pack_moves([r["moves"] for r in res], rl, ne)
    f = oracle.fills_from_moves(ne, rl - 5, words, engine.move_offsets(rl, ne))
    assert np.array_equal(f, [r["fills"] for r in res])
    for i in range(b.n_reads, len(seqs)):
        assert f[i] == int(ne[i]) * int(rl[i] - 5), i


def test_gradient_partials_vs_fd():
    """The partial derivatives in level_stdv (level_log_stdv held) and in
    level_log_stdv (level_stdv held) against central differences of the FP64
    score of the held path; their combination is the total derivative when
    level_log_stdv = log(level_stdv)."""
    from abea_amd import synth
    b = synth.workload("cpu64x2kb", seed=9, n=2)
    seq, ev, sc, sh, base, g, top = _grad_case(b, 0)
    g_read, g_mu, g_sg, g_lam = oracle.path_gradient(base["path"], base["from"], seq, ev, b.model,
                                                     sc, sh, partials=True)
    assert np.array_equal(g_read, g[0]) and np.array_equal(g_mu, g[1])
    np.testing.assert_allclose(g_sg + g_lam / b.model[:, 1].astype(np.float64), g[2],
                               rtol=1e-12, atol=1e-12)
    for r in top:
        n_cells = sum(1 for (k, _), f in zip(base["path"], base["from"])
                      if f != 2 and oracle.kmer_rank(seq[k:k + 6]) == r)
        assert g_lam[r] == -n_cells
        for col, an in ((1, g_sg[r]), (2, g_lam[r])):
            fs = []
            steps = []
            for sign in (1, -1):
                m = b.model.copy()
                v = np.float32(m[r, col] + sign * 1e-3)
                steps.append(float(v) - float(m[r, col]))
                m[r, col] = v
                fs.append(oracle.path_score(base["path"], base["from"], seq, ev, m, sc, sh))
            fd = (fs[0] - fs[1]) / (steps[0] - steps[1])
            assert abs(fd - an) <= 1e-6 * abs(an) + 1e-6, (r, col, fd, an)
