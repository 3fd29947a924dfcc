"""Replay f5c dump directories through libabea.

An f5c dump (GPU/dump_functions.h) holds, per batch, the inputs of
align_cuda() -- reads, event tables, scalings, the pore model -- and, when
f5c wrote them, its outputs (n_event_align_pairs, event_align_pairs).
replay_dump() runs every batch through the batched plan (event_t records
uploaded as they are) and compares each read's AlignedPair list with the
dump's; it can also write libabea's outputs in the same layout.
"""
from __future__ import annotations

import os

import numpy as np

from . import dump, engine
from ._lib import PLAN_BATCHED


def replay_dump(dump_dir: str, ctx: engine.Context | None = None, out_dir: str | None = None,
                form: int = PLAN_BATCHED) -> dict:
    """Returns {"batches", "reads", "compared", "mismatches": [(batch, read)]}."""
    own = ctx is None
    if own:
        ctx = engine.Context(0)
    report = {"batches": 0, "reads": 0, "compared": 0, "mismatches": []}
    try:
        nb = int(np.fromfile(os.path.join(dump_dir, "no_of_batches.dat"), "<i4", 1)[0])
        if out_dir:
            os.makedirs(out_dir, exist_ok=True)
            np.asarray([nb], "<i4").tofile(os.path.join(out_dir, "no_of_batches.dat"))
        for bi in range(nb):
            dbatch = dump.load_batch(os.path.join(dump_dir, str(bi)))
            b = dump.to_engine_batch(dbatch)
            plan = engine.Plan(ctx, engine.DeviceBatch.upload(b, ctx.device, aos_events=True),
                               form)
            plan.forward()
            res = plan.results()
            plan.close()
            got = []
            for i, rd in enumerate(dbatch.reads):
                off, n = int(res["pair_off"][i]), int(res["n_pairs"][i])
                pairs = res["pairs"][off:off + n]
                got.append(pairs)
                if rd.pairs is not None:
                    report["compared"] += 1
                    if not np.array_equal(pairs, rd.pairs):
                        report["mismatches"].append((bi, i))
            if out_dir:
                dump.write_outputs(os.path.join(out_dir, str(bi)),
                                   [g.shape[0] for g in got], got)
            report["batches"] += 1
            report["reads"] += len(dbatch.reads)
    finally:
        if own:
            ctx.close()
    return report
