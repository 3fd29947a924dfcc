"""Host-side check of the fast tiles' band-move step.

The core kernel's fast tiles (abea_kernels.hip: end_lanes, move_step,
band_fast) decide each band move on the scalar unit without recomputing the
band's lanes:

- Suzuki's rule ll < ur (GPU/align_cuda.cl:398-410), with both ends at -inf
  moving right on odd bands, is evaluated on the cells' bits as unsigned
  integers, reversed (bits(ll) > bits(ur)), the upper-right bits capped just
  below -inf's on odd bands.  This holds for cells <= 0 or -inf, which every
  cell of a read is when all its emissions are <= 0 (the staging checks it;
  other reads take the general band step);
- the lanes of the two end slots (K, K+99: the even one in the parity-0
  register at lane x0, the odd one in the parity-1 register at lane x1) and
  np = 1 - (K & 1) are scalar state: a right move maps (x0, x1) to
  (x1 + 1, x0) and flips the parity;
- the two slots masked to -inf after the band (K-1 and K+100 of the new K)
  are at lane x1 + np of the parity-0 register and lane x0 - np of the
  parity-1 register.

Lanes are used mod 64 (slots mod 128).  This restates those three steps and
checks them against the direct definitions: the float comparison, and the
lanes recomputed from K (slot k sits in lane (k mod 128) / 2 of register
k mod 2) over long random move sequences, including K < 0 (edge tiles start
at K >= -120).
"""
import numpy as np


def f32_bits(x):
    return int(np.array([x], np.float32).view(np.uint32)[0])


def move_by_bits(ll_bits: int, ur_bits: int, odd: bool) -> int:
    """The kernel's rule: right iff bits(ll) > bits(ur) (cap on odd bands)."""
    if odd:
        ur_bits = min(ur_bits, 0xFF7FFFFF)
    return int(ll_bits > ur_bits)


def move_by_floats(ll: np.float32, ur: np.float32, odd: bool) -> int:
    """The reference's rule."""
    if np.isneginf(ll) and np.isneginf(ur):
        return int(odd)
    return int(ll < ur)


def lanes_of(K: int):
    """(x0, x1, np) from K: lanes of the even and odd end slot, mod 64."""
    even, odd = (K, K + 99) if K % 2 == 0 else (K + 99, K)
    return (even % 128) // 2, (odd % 128) // 2, 1 - (K & 1)


def test_unsigned_bits_order_matches_float_compare():
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        -np.abs(rng.standard_normal(3000).astype(np.float32)) * np.float32(1e4),
        -np.abs(rng.standard_normal(1000).astype(np.float32)) * np.float32(1e-3),
        np.array([0.0, -np.inf, -1e-45, -3.4e38, -1.0, -1.0000001], np.float32),
    ]).astype(np.float32)
    pick = rng.integers(0, vals.shape[0], (20000, 2))
    # a share of exact ties, and of -inf on one or both ends
    pick[:2000, 1] = pick[:2000, 0]
    pick[2000:2500, 0] = vals.shape[0] - 5
    pick[2500:3000, 1] = vals.shape[0] - 5
    pick[3000:3010] = vals.shape[0] - 5  # both ends at -inf
    for a, b in pick:
        ll, ur = vals[a], vals[b]
        for odd in (False, True):
            assert move_by_bits(f32_bits(ll), f32_bits(ur), odd) == \
                move_by_floats(ll, ur, odd), (ll, ur, odd)


def test_positive_cells_break_the_bit_order():
    """Why reads with a positive emission take the general band step: above
    0 the bit order is the float order, not its reverse."""
    assert move_by_bits(f32_bits(np.float32(1.0)), f32_bits(np.float32(2.0)), False) == 0
    assert move_by_floats(np.float32(1.0), np.float32(2.0), False) == 1


def test_end_lanes_and_masks_follow_the_moves():
    rng = np.random.default_rng(7)
    for start in (-120, -51, -1, 0, 1, 37, 1000):
        for p_right in (0.1, 0.5, 0.9):
            K = start
            x0, x1, np_ = lanes_of(K)
            for r in (rng.random(4000) < p_right).astype(int):
                # the kernel's update
                if r:
                    x0, x1 = x1 + 1, x0
                    np_ ^= 1
                K += r
                # against the lanes recomputed from K
                e0, e1, enp = lanes_of(K)
                assert (x0 % 64, x1 % 64, np_) == (e0, e1, enp), (start, K)
                # masked slots K-1 and K+100: parity register and lane
                lo, hi = K - 1, K + 100
                slot0 = lo if lo % 2 == 0 else hi  # the even one, parity-0 register
                slot1 = hi if lo % 2 == 0 else lo  # the odd one, parity-1 register
                assert (x1 + np_) % 64 == (slot0 % 128) // 2, (start, K)
                assert (x0 - np_) % 64 == (slot1 % 128) // 2, (start, K)
