"""Host-side check of the band walk's staged-array indices.

The core kernel reads the staged k-mer records and event means without
bounds checks, relying on the layout the plan (and the allocation-free entry
point) gives them:

- each read's staged arrays carry STAGE_PAD = 128 padding entries on both
  sides;
- the arrays carry max(n_e + n_k) + 256 entries of slack past the last read,
  so a read whose unclamped prefetch runs past its own padding lands in a
  later read's region or in that slack (for slots whose cells are masked).

This test replays the walk's schedule on the host (abea_kernels.hip,
core_read: which bands run as interior or edge 8-band tiles, where runs
end and the slot registers are rebuilt, the general band's prefetches) for
band move sequences of real alignments (oracle walks of reads drawn from the
benchmark workloads, the refinement workload's perturbed 20 kb reads and
edge-case shapes) and for adversarial random move sequences, and checks every
index the kernel would read:

- 8-band tiles (their own operands and the next tile's): events in
  [E0 - 126, E0 + 16] and k-mers in [K0, K0 + 127] stay inside the read's
  own padding, [-128, n + 128);
- rebuilding the slots when a run ends: events inside the read's padding,
  k-mers at most 8 entries past it (into the next read's front padding);
- the general band's prefetch of event E+1 and k-mer K+128, and band 1's
  first prefetch: at most n_e + n_k + 51 and n_e + n_k + 77 (E <= n_e +
  n_k + 50 on any band), inside the slack.
"""
import numpy as np
import pytest

from abea_amd import synth

STAGE_PAD = 128
W = 100
HALF_W = 50


def walk_accesses(n_e: int, n_k: int, right: np.ndarray, e_guard: int = 111,
                  k_guard: int = 0) -> dict:
    """Min/max read-local index of each kind of staged-array read that
    core_read makes for one read, given its band moves (right[b] != 0 moves
    band b right of band b - 1; right[0], right[1] are ignored).  The
    schedule follows core_read / fast_run / tile_fast / restore_slots /
    band_general in abea_kernels.hip, whose tile guards are E0 <= n_e + 111
    (e_guard) and K0 <= n_k (k_guard = 0); the other values exist to show
    that the check fails when a guard is loosened."""
    nb = n_e + n_k + 2
    mv = np.asarray(right[:nb], np.int64).copy()
    mv[:2] = 0
    E = np.empty(nb, np.int64)
    K = np.empty(nb, np.int64)
    # band 1 at (E, K) = (50, -51); band b >= 2 from band b - 1
    steps_r = np.cumsum(mv[2:])
    K[1] = -1 - HALF_W
    E[1] = HALF_W
    K[2:] = K[1] + steps_r
    E[2:] = E[1] + (np.arange(1, nb - 1) - steps_r)
    acc = {}

    def note(kind, lo, hi):
        a = acc.get(kind)
        acc[kind] = (lo, hi) if a is None else (min(a[0], lo), max(a[1], hi))

    # band 1's prefetch: event E+1, k-mer K+128
    note("prefetch_ev", E[1] + 1, E[1] + 1)
    note("prefetch_km", K[1] + 128, K[1] + 128)

    def tile_kind(b):
        K0, E0 = K[b - 1], E[b - 1]
        if b % 8 or b + 8 > nb or K0 < -120 or K0 > n_k + k_guard or E0 < 0 or E0 > n_e + e_guard:
            return 0
        return 1 if (K0 >= 0 and K0 + 8 < n_k - W) else 2

    def more(bb, edge):
        K0, E0 = K[bb - 1], E[bb - 1]
        if bb + 8 > nb or E0 > n_e + e_guard:
            return False
        interior = K0 >= 0 and K0 + 8 < n_k - W
        return (K0 <= n_k + k_guard and not interior) if edge else interior

    def tile_reads(b0, ahead):
        # slot offsets o = (j - K0) mod 128 cover 0..127; event of slot o on
        # band b0 + i is b0 - 2 - K0 - o + i = E0 + 1 - o + i
        E0, K0 = E[b0 - 1], K[b0 - 1]
        i_hi = 15 if ahead else 7
        note("tile_ev", E0 + 1 - 127, E0 + 1 + i_hi)
        note("tile_km", K0, K0 + 127)

    b = 2
    n_tiles = [0, 0]
    while b < nb:
        kind = tile_kind(b)
        if kind:
            while kind:
                edge = kind == 2
                tile_reads(b, ahead=False)  # the run's first tile gathers its own
                while True:
                    tile_reads(b, ahead=True)  # and each tile the next one's
                    n_tiles[kind - 1] += 1
                    b += 8
                    if not more(b, edge):
                        break
                kind = tile_kind(b)
            if b >= nb:
                break
            Eb, Kb = E[b - 1], K[b - 1]
            note("restore_ev", Eb - 127, Eb)
            note("restore_km", Kb, Kb + 127)
            note("prefetch_ev", Eb + 1, Eb + 1)
            note("prefetch_km", Kb + 128, Kb + 128)
        # general band b: move, then prefetch for band b + 1
        note("prefetch_ev", E[b] + 1, E[b] + 1)
        note("prefetch_km", K[b] + 128, K[b] + 128)
        b += 1
    acc["tiles"] = tuple(n_tiles)
    acc["bands"] = nb
    return acc


def check_bounds(n_e: int, n_k: int, acc: dict):
    """The layout's guarantees for one read (read-local indices)."""
    slack = n_e + n_k + 256  # the plan's tail pad is max(n_e + n_k) + 256
    for kind, n in (("tile_ev", n_e), ("restore_ev", n_e), ("tile_km", n_k)):
        if kind in acc:
            lo, hi = acc[kind]
            assert lo >= -STAGE_PAD and hi < n + STAGE_PAD, (kind, lo, hi, n_e, n_k)
    if "restore_km" in acc:
        lo, hi = acc["restore_km"]
        assert lo >= -STAGE_PAD and hi <= n_k + 135, ("restore_km", lo, hi, n_e, n_k)
    lo, hi = acc["prefetch_ev"]
    assert lo >= -STAGE_PAD and hi <= n_e + n_k + 51 and hi < n_e + STAGE_PAD + slack
    lo, hi = acc["prefetch_km"]
    assert lo >= -STAGE_PAD and hi <= n_e + n_k + 77 and hi < n_k + STAGE_PAD + slack


def _oracle_moves(b, i, scale=None, shift=None):
    import oracle
    seq = b.read_seq(i)
    ev = b.ev_mean[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]]
    r = oracle.align(seq, ev, b.model, float(b.scale[i] if scale is None else scale[i]),
                     float(b.shift[i] if shift is None else shift[i]))
    return int(b.n_events[i]), max(int(b.read_len[i]) - 5, 0), r["moves"]


def test_fast_paths_cover_real_walks():
    """The schedule replay itself: on a real 8 kb walk most bands run in
    interior tiles, and the start takes edge tiles."""
    b = synth.generate([8000], seed=3)
    n_e, n_k, mv = _oracle_moves(b, 0)
    acc = walk_accesses(n_e, n_k, mv)
    check_bounds(n_e, n_k, acc)
    interior, edge = acc["tiles"]
    assert 8 * interior > 0.9 * acc["bands"], acc
    assert edge > 0, acc


def test_headline_workload_walks():
    """Reads with the benchmark workload's lengths (log-normal median 8 kb):
    a random sample and its longest reads, walked by the oracle."""
    lens = synth.workload_lengths("batched16384_ln8kb", seed=0)
    rng = np.random.default_rng(11)
    pick = np.concatenate([rng.choice(lens, 24, replace=False), np.sort(lens)[-2:]])
    b = synth.generate(pick, seed=5)
    for i in range(b.n_reads):
        n_e, n_k, mv = _oracle_moves(b, i)
        if n_e and n_k:
            check_bounds(n_e, n_k, walk_accesses(n_e, n_k, mv))


def test_refinement_workload_walks():
    """The refinement workload's 20 kb reads at perturbed scalings: their
    walks spend long stretches in the trim column and past the last event,
    where edge tiles run."""
    b = synth.generate([20000] * 6, seed=7)
    synth.perturb_scalings(b, seed=7)
    for i in range(b.n_reads):
        n_e, n_k, mv = _oracle_moves(b, i)
        check_bounds(n_e, n_k, walk_accesses(n_e, n_k, mv))


@pytest.mark.parametrize("shape", [(1, 1), (3, 200), (200, 3), (40, 2000), (2000, 40),
                                   (5000, 800), (800, 5000), (127, 128), (119, 1)])
@pytest.mark.parametrize("p_right", [0.0, 0.2, 0.5, 0.8, 1.0])
def test_random_move_sequences(shape, p_right):
    """Any move sequence, not only those a DP produces: the tile guards
    alone must keep the gathers in bounds."""
    n_e, n_k = shape
    nb = n_e + n_k + 2
    rng = np.random.default_rng(n_e * 7919 + n_k + int(p_right * 100))
    for trial in range(4):
        if trial < 2:
            mv = (rng.random(nb) < p_right).astype(np.uint8)
        else:
            # long runs of one move, then the other
            mv = np.zeros(nb, np.uint8)
            pos = 0
            while pos < nb:
                run = int(rng.integers(1, 400))
                mv[pos:pos + run] = rng.random() < p_right
                pos += run
        check_bounds(n_e, n_k, walk_accesses(n_e, n_k, mv))


@pytest.mark.parametrize("loosen", [dict(e_guard=127), dict(k_guard=8)])
def test_checker_catches_a_looser_guard(loosen):
    """The check is not vacuous: with either tile guard loosened (E0 <=
    n_e + 127 instead of n_e + 111, or K0 <= n_k + 8 instead of n_k), some
    adversarial walk's tile gathers leave the read's padding."""
    caught = 0
    for (n_e, n_k) in [(300, 400), (400, 300), (2000, 40), (40, 2000)]:
        for p_right in (0.0, 0.5, 1.0):
            mv = (np.random.default_rng(n_e + n_k).random(n_e + n_k + 2) < p_right)
            check_bounds(n_e, n_k, walk_accesses(n_e, n_k, mv.astype(np.uint8)))
            try:
                check_bounds(n_e, n_k, walk_accesses(n_e, n_k, mv.astype(np.uint8), **loosen))
            except AssertionError:
                caught += 1
    assert caught > 0
