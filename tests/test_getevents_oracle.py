"""Pinning the oracle's event detection (f5c getevents(), GPU/f5cmisc.h:44).

The body of getevents is f5c's port of scrappie's event detection; the
reference declares it and its event types (GPU/f5c.h:112-130) and holds no
raw signal, so there is nothing of the reference's to compare against
(parity with f5c unpinned).  These checks pin the restatement instead:
- closed forms: a piecewise-constant signal is cut exactly at its steps,
  with the event means its exact sums give, except that the detector's
  last boundary ends no event (scrappie's create_events counts one event
  per boundary, so the last event runs from the second-last boundary to
  the end); with noise every step is still a boundary;
- the windowed t-statistic against a direct float64 computation over the
  two windows (no prefix sums), and its zero edges;
- each event's mean and deviation against its own samples;
- hand-worked detector traces for each rule (rise above the running minimum,
  threshold, validity fall, w/2 age, masking of the long detector);
- planted mistakes in the oracle fail at least one of the above.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402


def step_signal(rng, n_seg, lo=7, hi=40, noise=0.0):
    """Piecewise-constant pA-like signal; consecutive levels differ by at
    least 5 pA.  Levels are multiples of 1/4, so that window sums, squares
    and means of a constant stretch are exact in float: on an arbitrary
    float level a window mean can miss the level by an ulp while the
    variance clamps to FLT_MIN, and the t-statistic then spikes inside a
    flat stretch (the algorithm's own behaviour on noise-free input).
    Returns (signal, boundaries, lengths, levels)."""
    lengths = rng.integers(lo, hi + 1, n_seg)
    grid = lambda: round(float(rng.uniform(60, 130)) * 4) / 4  # noqa: E731
    levels = [grid()]
    while len(levels) < n_seg:
        v = grid()
        if abs(v - levels[-1]) >= 5.0:
            levels.append(v)
    x = np.concatenate([np.full(L, np.float32(v), np.float32) for L, v in zip(lengths, levels)])
    if noise:
        x = (x + rng.normal(0, noise, x.shape[0])).astype(np.float32)
    bounds = np.cumsum(lengths)[:-1]
    return x, bounds.astype(np.int64), lengths, np.array(levels, np.float32)


def expected_events(lengths, levels):
    """Events of a clean step signal: one per segment, the last two segments
    merged (the detector's last boundary, the last step, ends no event).
    Returns (starts, lengths, means): a segment's double sum L*v is exact,
    and an event mean is its sum rounded to float, then divided by its
    length in float."""
    sums = lengths * levels.astype(np.float64)
    lens = np.concatenate([lengths[:-2], [lengths[-2] + lengths[-1]]]) if len(lengths) > 1 \
        else lengths.copy()
    tot = np.concatenate([sums[:-2], [sums[-2] + sums[-1]]]) if len(lengths) > 1 else sums
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    means = (np.float32(tot) / lens.astype(np.float32)).astype(np.float32)
    return starts, lens, means


def test_step_signals_cut_exactly_at_steps():
    rng = np.random.default_rng(11)
    for _ in range(20):
        x, bounds, lengths, levels = step_signal(rng, int(rng.integers(2, 60)))
        S, Q = oracle.signal_prefix(x)
        assert oracle.detect_peaks(oracle.tstat(S, Q, 3), oracle.tstat(S, Q, 6)).tolist() == \
            bounds.tolist()
        ev = oracle.getevents(x)
        starts, lens, means = expected_events(lengths, levels)
        assert ev.shape[0] == lengths.shape[0] - 1
        assert np.array_equal(ev["start"].astype(np.int64), starts)
        assert np.array_equal(ev["length"], lens.astype(np.float32))
        assert np.array_equal(ev["mean"], means)
        # the variance is a difference of rounded terms: at most a few ulps
        # of v^2 on one level; the merged last event holds two levels
        assert np.all(ev["stdv"][:-1] <= np.sqrt(8 * np.spacing(levels[:-2] * levels[:-2])))
        a, b = levels[-2].astype(np.float64), levels[-1].astype(np.float64)
        L1, L2 = float(lengths[-2]), float(lengths[-1])
        sd = abs(a - b) * np.sqrt(L1 * L2) / (L1 + L2)
        assert abs(float(ev["stdv"][-1]) - sd) <= 1e-3 * sd + 1e-2


def test_noisy_steps_are_cut_and_stats_match_samples():
    # noise over-segments flat stretches (a 3-sample t-statistic of pure
    # noise often passes 1.4), but every step, whose statistic dwarfs its
    # neighbours', stays a cut
    rng = np.random.default_rng(12)
    for _ in range(10):
        x, bounds, lengths, _ = step_signal(rng, int(rng.integers(5, 80)), noise=0.05)
        S, Q = oracle.signal_prefix(x)
        peaks = oracle.detect_peaks(oracle.tstat(S, Q, 3), oracle.tstat(S, Q, 6))
        assert set(bounds.tolist()) <= set(peaks.tolist())
        ev = oracle.getevents(x)
        starts = ev["start"].astype(np.int64)
        assert np.array_equal(starts, np.concatenate([[0], peaks[:-1]]))
        assert np.all(np.diff(starts) > 0)
        assert int(ev["start"][-1] + ev["length"][-1]) == x.shape[0]
        for e in ev:
            seg = x[int(e["start"]):int(e["start"]) + int(e["length"])].astype(np.float64)
            assert abs(e["mean"] - seg.mean()) <= 1e-5 * abs(seg.mean())
            # the variance is E[x^2] - mean^2 in float: exact up to a few ulps
            # of mean^2 (cancellation), whatever the deviation itself
            ulp = float(np.spacing(np.float32(seg.mean() ** 2)))
            assert abs(float(e["stdv"]) ** 2 - seg.var()) <= 6 * ulp


def _tstat_direct(x, w):
    """The statistic in float64 from the two windows, and a bound on the
    float restatement's error: its variance carries a few ulps of the
    squared means (E[x^2] - m^2 cancels) and its means an ulp each, so
    |dt|/t <= 4 ulp(m^2)/v + 2 ulp(m)/|dm| (to first order)."""
    x = x.astype(np.float64)
    n = x.shape[0]
    t = np.zeros(n)
    tol = np.zeros(n)
    if n < 2 * w or w < 2:
        return t, tol
    for i in range(w, n - w + 1):
        a, b = x[i - w:i], x[i:i + w]
        v = max(a.var() + b.var(), np.finfo(np.float32).tiny)
        dm = abs(b.mean() - a.mean())
        t[i] = dm / np.sqrt(v / w)
        m = max(abs(a.mean()), abs(b.mean()))
        tol[i] = t[i] * (4 * float(np.spacing(np.float32(m * m))) / v
                         + 2 * float(np.spacing(np.float32(m))) / max(dm, 1e-30) + 1e-5)
    return t, tol


def test_tstat_against_direct_windows():
    rng = np.random.default_rng(13)
    for n in (5, 11, 12, 13, 300):
        x, _, _, _ = step_signal(rng, 40, noise=1.5)
        x = x[:n]
        S, Q = oracle.signal_prefix(x)
        assert S[-1] == np.sum(x.astype(np.float64))  # exact here (pA-scale floats)
        for w in (3, 6):
            t = oracle.tstat(S, Q, w)
            ref, tol = _tstat_direct(x, w)
            if n < 2 * w:
                assert np.all(t == 0)
                continue
            assert np.all(t[:w] == 0) and np.all(t[n - w + 1:] == 0)
            m = ref > 1e-2
            assert np.all(np.abs(t[m] - ref[m]) <= tol[m])


def test_detector_rules_hand_worked():
    z = np.zeros(40, np.float32)
    # short peak of 5 at 10; falls to 4 at 11 (valid); emitted once 2 old
    ts = z.copy()
    ts[9:13] = [1.0, 5.0, 4.0, 0.0]
    assert oracle.detect_peaks(ts, z).tolist() == [10]
    # below the short threshold: never valid
    ts = z.copy()
    ts[9:13] = [0.5, 1.3, 0.0, 0.0]
    assert oracle.detect_peaks(ts, z).tolist() == []
    # a rise of no more than the peak height (0.2) above the minimum is no
    # peak; one of more is (2.25 - 2 = 0.25, then a fall and an age of 2)
    ts = np.full(40, 2.0, np.float32)
    ts[10] = 2.125
    assert oracle.detect_peaks(ts, z).tolist() == []
    ts[10] = 2.25
    ts[11:] = 0.0
    assert oracle.detect_peaks(ts, z).tolist() == [10]
    # after the validity fall, a higher value within w/2 moves the peak
    ts = z.copy()
    ts[9:14] = [0.0, 5.0, 4.7, 6.0, 0.0]
    assert oracle.detect_peaks(ts, z).tolist() == [12]
    # the long detector alone (threshold 9, emitted when more than 3 old)
    tl = z.copy()
    tl[19:23] = [1.0, 12.0, 3.0, 0.0]
    assert oracle.detect_peaks(z, tl).tolist() == [20]
    # ... below its threshold: nothing
    tl[20] = 8.5
    assert oracle.detect_peaks(z, tl).tolist() == []
    # a short peak above threshold at 20 masks the long detector through 23
    # and restarts it: the long peak at 22 is lost, the short one stands
    ts = z.copy()
    ts[19:23] = [0.0, 3.0, 0.0, 0.0]
    tl = z.copy()
    tl[21:25] = [1.0, 12.0, 3.0, 0.0]
    assert oracle.detect_peaks(ts, tl).tolist() == [20]
    # a long peak after the mask is its own boundary, in emission order
    tl = z.copy()
    tl[29:33] = [1.0, 12.0, 3.0, 0.0]
    assert oracle.detect_peaks(ts, tl).tolist() == [20, 30]


def test_edge_signals():
    # shorter than the short window pair: one event over the whole signal
    for n in (1, 2, 5):
        x = np.arange(n, dtype=np.float32) + 70
        ev = oracle.getevents(x)
        assert ev.shape[0] == 1 and ev["start"][0] == 0 and ev["length"][0] == n
        assert np.float32(ev["mean"][0]) == np.float32(np.float32(x.astype(np.float64).sum()) / n)
    # a constant signal: one event
    ev = oracle.getevents(np.full(1000, 88.5, np.float32))
    assert ev.shape[0] == 1 and ev["mean"][0] == np.float32(88.5) and ev["stdv"][0] == 0


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


def _step_means(lengths, levels):
    return expected_events(lengths, levels)[2]


def _long_step_signal():
    # long enough that float prefix sums lose the exact means
    rng = np.random.default_rng(14)
    return step_signal(rng, 3000, lo=20, hi=40)


@pytest.mark.parametrize("mutant", ["ORC_MUTATE_GE_FLOAT_PREFIX", "ORC_MUTATE_GE_NO_MASK",
                                    "ORC_MUTATE_GE_EMIT_EARLY",
                                    "ORC_MUTATE_GE_KEEP_LAST_BOUNDARY"])
def test_getevents_checks_catch_planted_mistakes(mutant):
    so = _mutant(mutant)

    def run():
        failures = 0
        try:
            test_detector_rules_hand_worked()
        except AssertionError:
            failures += 1
        x, bounds, lengths, levels = _long_step_signal()
        ev = oracle.getevents(x)
        if not (ev.shape[0] == lengths.shape[0] - 1 and np.array_equal(ev["mean"], _step_means(
                lengths, levels))):
            failures += 1
        return failures

    assert _with_lib(so, run) > 0
    # and the real oracle passes the same checks
    x, bounds, lengths, levels = _long_step_signal()
    ev = oracle.getevents(x)
    assert np.array_equal(ev["mean"], _step_means(lengths, levels))
