"""Loader for the repacked golden f5c dump subset (tests/golden/*.npz)."""
from __future__ import annotations

import numpy as np

from .synth import Batch


def load_golden_subset(path: str):
    """Returns (Batch, golden_pairs list, raw npz dict)."""
    z = dict(np.load(path, allow_pickle=False))
    n = z["seq_len"].shape[0]
    ev_ptr = z["ev_ptr"][:-1].astype(np.int64)
    n_events = np.diff(z["ev_ptr"]).astype(np.int32)
    sc = z["scalings"]
    b = Batch(model=z["model"].astype(np.float32), seq=z["seq"].astype(np.uint8),
              seq_ptr=z["seq_ptr"][:-1].astype(np.int64), read_len=z["seq_len"].astype(np.int32),
              ev_mean=z["ev_mean"].astype(np.float32), ev_ptr=ev_ptr, n_events=n_events,
              scale=sc[:, 0].copy(), shift=sc[:, 1].copy(), var=sc[:, 2].copy(),
              log_var=sc[:, 3].copy(), ev_stdv=z["ev_stdv"], ev_length=z["ev_length"],
              ev_start=z["ev_start"])
    pp = z["pair_ptr"]
    golden = [z["pairs"][pp[i]:pp[i + 1]] for i in range(n)]
    return b, golden, z
