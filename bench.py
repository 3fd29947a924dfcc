"""ABEA benchmark on one MI355X.

Headline: batched persistent ABEA (length-sorted queue, persistent core that
stages, walks and traces back each read, replayed as a hipGraph, autotuned
launch shapes) over the 16,384-read long-tailed workload (log-normal median
8 kb, clipped 500 b-200 kb, ~150 Mbases, ~40 G nominal cells), inputs
resident in HBM as f5c event_t records.  A "step" is one full forward over
that batch; the core kernel's time comes from hipEvents recorded inside the
same graph steps.

Also reported (same JSON line):
- the plain per-read-launch form against the batched form on all 4,096 reads
  of the log-normal (median 5 kb) workload, and on the 64 x 2 kb reads;
- the host-buffer drop-ins: align_cuda()'s batch entry on the 4,096 reads
  with its phases (flatten, H2D, kernels, D2H, unflatten) and align() per
  read, sequential and from 8 threads;
- the adjoint/forward time ratio on 1,024 x 10 kb reads;
- postalign/recalibrate/MOM and getevents on their workloads;
- the scaling-refinement loop on 8,192 x 20 kb reads;
- the CPU leg: the oracle (f5c semantics, OpenMP over reads) timed on a
  bounded sample of the headline workload with all host threads and with
  one, and on the 64 x 2 kb reads; every read of the headline batch (the
  timed sample, plus the 64 longest, then the rest untimed) compared read by
  read with the GPU's headline results; the FP64
  error of the float scores; fills; finite differences of the adjoint.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "abea-opencl_amd"))

HBM_PEAK_GBS = 8000.0
W = 100
# Per launch of the core kernel on the headline workload, from rocprofv3
# (profiles/abea_kernels_batched16384_ln8kb.md gives the commands): VALU
# wave-instructions (SQ_INSTS_VALU), which depend on the traceback overlap
# only, and HBM bytes (FETCH_SIZE + WRITE_SIZE passes, scaled by a 1 GiB
# copy's readings) per launch shape (kernel, waves per CU, traceback overlap).
# The autotuner's unfused candidates (a separate post kernel) have no entry.
PROFILED_VALU = {  # (kernel, traceback overlap)
    ("abea_core_kernel<1, true>", 128): 2.05535e10,
    ("abea_core_kernel<2, true>", 128): 2.05537e10,
    ("abea_core_kernel<4, true>", 128): 2.05537e10,
    ("abea_core_kernel<1, true>", 64): 2.06509e10,
    ("abea_core_kernel<1, true>", 256): 2.06105e10,
}
PROFILED_TRAFFIC = {  # (kernel, waves per CU, traceback overlap)
    ("abea_core_kernel<1, true>", 16, 128): 46036583104,
    ("abea_core_kernel<1, true>", 20, 128): 48777037088,
    ("abea_core_kernel<1, true>", 24, 128): 48763236352,
    ("abea_core_kernel<1, true>", 32, 128): 48762137568,
    ("abea_core_kernel<2, true>", 32, 128): 48772339919,
    ("abea_core_kernel<4, true>", 32, 128): 48716171416,
    ("abea_core_kernel<1, true>", 24, 64): 48962427589,
    ("abea_core_kernel<1, true>", 24, 256): 51101081867,
}
N_SIMD = 1024          # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9       # peak engine clock


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(b, path_len) -> tuple[int, int]:
    """Algorithmic HBM bytes of one forward, split (band walk, traceback).

    Per read: W*n_bands (a one-byte from-code per band cell) + 4*n_e (event
    means) + 2*n_k (k-mer ranks) + n_bands/8 (band-move bits) + 16
    (scalings) for the band walk; path_len (from-codes read back) +
    8*path_len (aligned pairs) for the traceback.  The library stores 2-bit
    codes, so its actual traffic is below this figure; the pmc summaries
    under profiles/ give the measured bytes."""
    nb = b.n_bands.astype("int64")
    ok = (b.n_kmers > 0) & (b.n_events > 0)
    walk = int((W * nb[ok]).sum() + 4 * b.n_events[ok].sum() + 2 * b.n_kmers[ok].sum()
               + (nb[ok] // 8).sum() + 16 * ok.sum())
    tb = int(9 * path_len[ok].astype("int64").sum())
    return walk, tb


def kernel_name(tune) -> str:
    return (f"abea_core_kernel<{tune['core_waves_per_block']}, "
            f"{'true' if tune['fuse_traceback'] else 'false'}>")


def timed_steps(plan, steps, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.forward()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def best_ms(fn, torch, reps=3):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return 1e3 * min(ts)


def hbm_copy_gbs(torch, gib=2, reps=5):
    """The box's own HBM rate: a device-to-device copy of `gib` GiB (read
    once, written once), best of `reps`, hipEvent-timed on torch's stream.
    The vendor figure (8 TB/s) stays the roofline's peak; this says what a
    plain copy reaches on the same box."""
    n = (gib << 30) // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = 1e30
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    torch.cuda.empty_cache()
    return round(2 * (gib << 30) / (best * 1e-3) / 1e9, 1)


# --------------------------------------------------------------------- legs
def headline_leg(args, ctx):
    import numpy as np
    import torch

    from abea_amd import engine, synth
    from abea_amd._lib import PLAN_BATCHED

    t = time.perf_counter()
    wl = "batched16384_ln8kb"
    b = synth.workload(wl, seed=args.seed, with_aos=True)
    cells = b.nominal_cells()
    log(f"[bench] generated {wl}: {b.n_reads} reads, {int(b.read_len.sum()) / 1e6:.1f} Mb, "
        f"{int(b.n_events.sum()) / 1e6:.1f} M events, {cells / 1e9:.2f} G cells "
        f"in {time.perf_counter() - t:.1f}s")
    db = engine.DeviceBatch.upload(b, 0, aos_events=True)
    plan = engine.Plan(ctx, db, PLAN_BATCHED)
    log(f"[bench] plan workspace {plan.workspace_bytes() / 1e9:.1f} GB")
    plan.forward()
    torch.cuda.synchronize()
    if args.profile:
        # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE
        # readings scale the forward's counters to bytes
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 1 GiB read + 1 GiB written")
        del src, dst
    if args.tuning:
        wpb, wpc, fuse, ovl = (int(x) for x in args.tuning.split(","))
        plan.set_tuning(core_waves_per_block=wpb, core_waves_per_cu=wpc, fuse_traceback=fuse,
                        traceback_overlap=ovl)
    elif not args.profile:
        t = time.perf_counter()
        # two forwards per candidate: the shapes differ by 1-3%, about the
        # noise of a single forward
        plan.autotune(reps=2)
        log(f"[bench] autotuned in {time.perf_counter() - t:.1f}s")
    tune = plan.tuning()
    log(f"[bench] tuning {tune}")

    # timed steps: graph replays, each timed on the plan's stream by hipEvents
    # around its launch (the graph holds the one-wave queue reset and the
    # core); eager steps (--eager) time each kernel
    plan.set_tuning(use_graph=0 if args.eager else 1)
    for _ in range(args.warmup):
        plan.forward()
    torch.cuda.synchronize()
    kt = np.zeros(3)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.forward()
        kt += np.array(plan.times()[:3])
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    kt /= args.steps
    if args.eager:
        kt_source = "hipEvents between the kernels of each timed eager step"
    else:
        kt_source = ("hipEvents around each timed graph replay on the plan's stream: the core "
                     "plus the one-wave queue reset before it")
        kt[0] = kt[2] = -1.0
    gcups = cells / sec / 1e9
    log(f"[bench] batched {'eager' if args.eager else 'graph'}: {sec * 1e3:.2f} ms/step, "
        f"{gcups:.1f} GCUPS; kernel ms queue_reset/core/post {kt.round(3).tolist()}")
    res = plan.results()
    moves = plan.moves()
    qc_pass = int((res["n_pairs"] > 0).sum())
    walk_b, tb_b = alg_bytes(b, res["path_len"])
    # the dominant kernel is the core; fused, it also does the traceback
    core_bytes = walk_b + (tb_b if tune["fuse_traceback"] else 0)
    achieved = core_bytes / (kt[1] * 1e-3) / 1e9
    kname = kernel_name(tune)
    traffic = PROFILED_TRAFFIC.get((kname, tune["core_waves_per_cu"], tune["traceback_overlap"]))
    valu = PROFILED_VALU.get((kname, tune["traceback_overlap"]))
    out = {
        "metric": "ABEA GCUPS (batched persistent form, 1 GPU)",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cells, f64 transition sums",
        "data": "synthetic (seeded; f5c event_t records in HBM)",
        "config": {"workload": "batched16384_ln8kb", "reads": b.n_reads,
                   "bases": int(b.read_len.sum()), "events": int(b.n_events.sum()),
                   "nominal_cells": cells, "bandwidth": W, "tuning": tune},
        "reads_per_s": round(b.n_reads / sec, 1),
        "kernel_ms": {"queue_reset": round(kt[0], 3) if kt[0] >= 0 else None,
                      "core": round(kt[1], 3),
                      "post": round(kt[2], 3) if kt[2] >= 0 else None, "source": kt_source},
        "qc_pass_reads": qc_pass,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "alg_bytes_per_launch": core_bytes,
                     "box_copy_gbs": None if args.profile else hbm_copy_gbs(torch)},
    }
    if valu:
        # VALU issue rate against the ceiling of one wave instruction per SIMD
        # per 4 cycles at the peak clock
        rate = valu / (kt[1] * 1e-3)
        peak = N_SIMD * CLOCK_HZ / 4
        out["issue"] = {"bound": "valu", "valu_insts_per_launch": valu,
                        "valu_insts_per_band": round(valu / int(b.n_bands.sum()), 2),
                        "achieved": round(rate / 1e9, 1), "peak": round(peak / 1e9, 1),
                        "unit": "G wave-instructions/s", "frac": round(rate / peak, 4)}
    if traffic is None:
        out["roofline"]["traffic_note"] = "no pmc profile at this launch shape"
    keep = {"pairs": res["pairs"], "path_from": res["path_from"], "path_len": res["path_len"],
            "n_pairs": res["n_pairs"], "score": res["score"], "avg": res["avg"],
            "pair_off": res["pair_off"], "moves": moves}
    del plan, db, res
    torch.cuda.empty_cache()
    return out, b, keep, tune


def plain_leg(args, ctx, tune):
    """The reference's shape (one launch set per read) against the batched
    form on the same reads: all 4,096 reads of the 5 kb log-normal workload,
    and the 64 x 2 kb reads."""
    import numpy as np  # noqa: F401
    import torch

    from abea_amd import engine, synth
    from abea_amd._lib import PLAN_BATCHED, PLAN_PLAIN

    res = {}
    for wl, b1 in (("plain4096_ln5kb", synth.workload("plain4096_ln5kb", seed=args.seed + 1)),
                   ("cpu64x2kb", synth.workload("cpu64x2kb", seed=args.seed + 2))):
        d1 = engine.DeviceBatch.upload(b1, 0)
        p_plain = engine.Plan(ctx, d1, PLAN_PLAIN)
        p_plain.forward()
        s_plain = timed_steps(p_plain, 1, torch)
        p_b = engine.Plan(ctx, d1, PLAN_BATCHED)
        p_b.set_tuning(**{k: v for k, v in tune.items()})
        p_b.forward()
        s_b = timed_steps(p_b, 5, torch)
        cells = b1.nominal_cells()
        res[wl] = {"reads": b1.n_reads, "nominal_cells": cells,
                   "plain_gcups": round(cells / s_plain / 1e9, 3),
                   "plain_reads_per_s": round(b1.n_reads / s_plain, 1),
                   "batched_gcups": round(cells / s_b / 1e9, 3),
                   "batched_reads_per_s": round(b1.n_reads / s_b, 1),
                   "batched_over_plain": round(s_plain / s_b, 1)}
        log(f"[bench] plain/batched {wl}: {res[wl]}")
        del p_plain, p_b, d1
    return res


def dropin_leg(args):
    """Host buffers in and out, PCIe included: align_cuda()'s batch entry
    (abea_align_db) on the 4,096-read workload, with the phases the
    reference's host times, and align() (abea_align) per read."""
    import numpy as np

    from abea_amd import _lib, synth

    L = _lib.lib()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free.restype = None
    def host_batch(b):
        n = b.n_reads
        aos = b.events_aos()
        reads, ets = (C.c_char_p * n)(), (_lib.EventTable * n)()
        seqs = [b.read_seq(i) for i in range(n)]
        base = aos.ctypes.data
        for i in range(n):
            reads[i] = seqs[i]
            ets[i] = _lib.EventTable(int(b.n_events[i]), 0, int(b.n_events[i]),
                                     C.cast(base + 24 * int(b.ev_ptr[i]), C.POINTER(_lib.EventT)))
        return aos, seqs, reads, ets

    def align_db(wl, b):
        n = b.n_reads
        model = np.ascontiguousarray(b.model, np.float32)
        aos, seqs, reads, ets = host_batch(b)
        rl = np.ascontiguousarray(b.read_len, np.int32)
        scal = np.ascontiguousarray(b.scalings(), np.float32)
        outp = (C.POINTER(_lib.AlignedPair) * n)()
        nout = np.zeros(n, np.int32)
        phases = []
        walls = []
        for rep in range(3):
            t0 = time.perf_counter()
            st = L.abea_align_db(n, C.cast(reads, C.c_void_p), rl.ctypes.data_as(C.c_void_p),
                                 C.cast(ets, C.c_void_p), model.ctypes.data_as(C.c_void_p),
                                 scal.ctypes.data_as(C.c_void_p), C.cast(outp, C.c_void_p),
                                 nout.ctypes.data_as(C.c_void_p))
            walls.append(time.perf_counter() - t0)
            _lib.check(st, "abea_align_db")
            t = (C.c_double * 7)()
            L.abea_host_last_times(t)
            phases.append(list(t))
            for i in range(n):
                libc.free(C.cast(outp[i], C.c_void_p))
        k = int(np.argmin(walls[1:])) + 1  # the first call also sizes the pooled buffers
        ph = phases[k]
        # bytes of the input copies and of the output copies (the counts,
        # then the pairs, compacted on the device).  The bases and event
        # means go to the device during the flatten, by the flattening
        # threads; the "h2d" phase holds the rest (scalings, model, CSR
        # offsets and counts)
        staged_bytes = int(b.read_len.sum()) + 4 * int(b.n_events.sum())
        rest_bytes = 16 * n + 12 * 4096 + 24 * n
        h2d_bytes = staged_bytes + rest_bytes
        n_pairs = int(nout.astype(np.int64).sum())
        d2h_bytes = 4 * n + 8 * n_pairs
        cells = b.nominal_cells()
        return {"workload": wl, "reads": n, "qc_pass_reads": int((nout > 0).sum()),
                "gcups": round(cells / walls[k] / 1e9, 3),
                "reads_per_s": round(n / walls[k], 1),
                "ms": {"flatten": round(ph[0], 3), "h2d": round(ph[1], 3),
                       "plan": round(ph[2], 3), "kernels": round(ph[3], 3),
                       "d2h": round(ph[4], 3), "unflatten": round(ph[5], 3),
                       "total": round(ph[6], 3)},
                "h2d_bytes": h2d_bytes, "h2d_bytes_during_flatten": staged_bytes,
                "d2h_bytes": d2h_bytes,
                "d2h_gbs": round(d2h_bytes / (ph[4] * 1e-3) / 1e9, 2),
                "pairs_returned": n_pairs}

    b = synth.workload("plain4096_ln5kb", seed=args.seed + 1, with_aos=True)
    res = {"align_db": align_db("plain4096_ln5kb", b)}
    # the same lengths at the real reads' k-mer skip rate: every read passes
    # QC and returns its pairs
    bq = synth.generate(b.read_len, seed=args.seed + 1, with_aos=True, skip=0.015)
    res["align_db_qc_pass"] = align_db("plain4096_ln5kb, skip rate 1.5%", bq)
    del bq
    log(f"[bench] align_db: {res}")
    n = b.n_reads
    model = np.ascontiguousarray(b.model, np.float32)
    aos, seqs, reads, ets = host_batch(b)
    # align() per read: the first 64 reads, one thread, then 8 threads at once
    m = 64
    outs = [np.zeros((2 * int(b.n_events[i]) + 1, 2), np.int32) for i in range(m)]

    def one(i):
        ev = ets[i]
        sc = _lib.ScalingsT(float(b.scale[i]), float(b.shift[i]), 0.0, 0.0)
        return L.abea_align(outs[i].ctypes.data_as(C.c_void_p), seqs[i], len(seqs[i]), ev,
                            model.ctypes.data_as(C.c_void_p), sc, 4000.0)

    for i in range(8):
        one(i)
    t0 = time.perf_counter()
    ph1 = np.zeros(7)
    for i in range(m):
        one(i)
        t = (C.c_double * 7)()
        L.abea_host_last_times(t)
        ph1 += np.array(list(t))
    seq_s = time.perf_counter() - t0
    ph1 /= m

    def worker(t):
        for i in range(t, m, 8):
            one(i)

    def threaded():
        ts = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
        t0 = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return time.perf_counter() - t0

    threaded()  # creates and sizes the pooled contexts the threads use
    thr_s = threaded()
    cells_m = int(W * b.n_bands[:m].sum())
    res["align_per_read"] = {"reads": m, "one_thread_reads_per_s": round(m / seq_s, 1),
                             "one_thread_gcups": round(cells_m / seq_s / 1e9, 3),
                             "eight_threads_reads_per_s": round(m / thr_s, 1),
                             "ms_per_read": round(1e3 * seq_s / m, 3),
                             "ms_per_read_phases": dict(zip(
                                 ["flatten", "h2d", "plan", "kernels", "d2h", "unflatten",
                                  "total"], [round(float(x), 3) for x in ph1])),
                             "host_contexts": int(L.abea_host_contexts())}
    log(f"[bench] drop-ins: {res}")
    return res


def adjoint_leg(args, ctx, tune):
    import numpy as np
    import torch

    from abea_amd import engine, synth
    from abea_amd._lib import PLAN_BATCHED

    b3 = synth.workload("adjoint1024x10kb", seed=args.seed + 3)
    d3 = engine.DeviceBatch.upload(b3, 0)
    p3 = engine.Plan(ctx, d3, PLAN_BATCHED)
    p3.set_tuning(**{k: v for k, v in tune.items()})
    p3.forward()
    s_fwd = timed_steps(p3, 3, torch)
    gsc = torch.ones(b3.n_reads, device="cuda")
    p3.adjoint(gsc)
    torch.cuda.synchronize()
    t = time.perf_counter()
    dev_ms = []
    for _ in range(3):
        p3.adjoint(gsc)
        dev_ms.append(p3.times()[3])  # clear + adjoint + finish kernels, hipEvents
    torch.cuda.synchronize()
    s_adj = (time.perf_counter() - t) / 3
    # two adjoints of the same forward give the same bits (fixed-point sums)
    g1 = p3.adjoint(gsc)[1].cpu().numpy()
    g2 = p3.adjoint(gsc)[1].cpu().numpy()
    res = {"workload": "adjoint1024x10kb", "forward_ms": round(s_fwd * 1e3, 3),
           "adjoint_ms": round(s_adj * 1e3, 3),
           "adjoint_kernels_ms": round(float(np.mean(dev_ms)), 3),
           "adjoint_over_forward": round(s_adj / s_fwd, 4),
           "deterministic": bool(np.array_equal(g1, g2)),
           "note": "adjoint_ms is the call's wall time (outputs allocated, one sync); "
                   "adjoint_kernels_ms its kernels on the plan's stream"}
    log(f"[bench] adjoint: {res}")
    del p3, d3
    # eight of the reads for the finite-difference check of the CPU leg
    sub = b3.subset(np.arange(8))
    ps = engine.Plan(ctx, engine.DeviceBatch.upload(sub, 0), PLAN_BATCHED)
    ps.forward()
    w = torch.ones(sub.n_reads, device="cuda")
    g_scal, g_model = ps.adjoint(w)
    fd_in = {"batch": sub, "res": ps.results(), "moves": ps.moves(),
             "g_scal": g_scal.cpu().numpy().astype(np.float64),
             "g_model": g_model.cpu().numpy()}
    ps.close()
    return res, fd_in


def postalign_leg(args, ctx, tune) -> dict:
    """f5c's steps around align() on 16,384 reads that pass align()'s QC (log-
    normal median 8 kb, the real-data k-mer skip rate): MOM scalings, forward,
    then postalign + recalibrate_model, with and without the event alignment
    records."""
    import numpy as np
    import torch

    from abea_amd import engine, synth
    from abea_amd._lib import PLAN_BATCHED

    b5 = synth.workload("postalign16384_ln8kb", seed=args.seed + 5)
    d5 = engine.DeviceBatch.upload(b5, 0)
    p5 = engine.Plan(ctx, d5, PLAN_BATCHED)
    if tune:
        p5.set_tuning(**{k: v for k, v in tune.items()})
    p5.forward()
    torch.cuda.synchronize()
    fwd_ms = best_ms(p5.forward, torch)
    mom_ms = best_ms(lambda: p5.estimate_scalings(), torch)
    full_ms = best_ms(lambda: p5.postalign(), torch)
    nrec_ms = best_ms(lambda: p5.postalign(records=False), torch)
    pr = engine.postalign_results(p5, p5.postalign())
    n_pairs = int(p5.n_pairs[:b5.n_reads].sum().item())
    n_rec = int(pr["n_alignment"].sum())
    n_km = int(np.maximum(b5.read_len.astype(np.int64) - 5, 0).sum())
    # algorithmic bytes: pairs read (8 B each), records written (32 B),
    # map entries written (8 B per k-mer), an event mean read (4 B) per
    # record (an upper bound: only 'M' records need one); model rows are
    # cache-resident
    post_bytes = 8 * n_pairs + 32 * n_rec + 8 * n_km + 4 * n_rec
    flags = np.bincount(pr["read_stat_flag"], minlength=5)
    res = {
        "workload": "postalign16384_ln8kb", "reads": b5.n_reads, "pairs": n_pairs,
        "records": n_rec, "forward_ms": round(fwd_ms, 3),
        "estimate_scalings_ms": round(mom_ms, 3),
        "postalign_recalibrate_ms": round(full_ms, 3),
        "postalign_recalibrate_no_records_ms": round(nrec_ms, 3),
        "alg_bytes": post_bytes,
        "achieved_gbs": round(post_bytes / (full_ms * 1e-3) / 1e9, 1),
        "pairs_per_s": round(n_pairs / (full_ms * 1e-3), 1),
        "status": {"ok": int(flags[0]), "failed_calibration": int(flags[1]),
                   "failed_alignment": int(flags[2]), "failed_quality_chk": int(flags[4])}}
    log(f"[bench] postalign: {res}")
    del p5, d5, b5, pr
    return res


def getevents_leg(args, ctx):
    """f5c's step before MOM and align(): event detection on the raw signals
    (pA) of 1,024 reads of the 5 kb log-normal workload, resident in HBM.
    Returns the leg and the host signal for the CPU leg's timing."""
    import torch

    from abea_amd import _lib, engine, synth

    b = synth.workload("plain4096_ln5kb", seed=args.seed + 7, n=1024)
    raw_h, raw_off = synth.raw_signal(b, seed=args.seed + 7)
    raw = torch.from_numpy(raw_h).cuda()
    n_samples = int(raw_off[-1])
    ev = torch.empty((n_samples + b.n_reads) * 24, dtype=torch.uint8, device="cuda")
    cnt = torch.empty(b.n_reads, dtype=torch.int64, device="cuda")
    engine.getevents(ctx, raw, raw_off, ev, cnt)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(ctx.stream)
        _lib.check(_lib.lib().abea_getevents_device(ctx.handle, b.n_reads,
                                                     raw_off.ctypes.data_as(_lib.P),
                                                     _lib.P(raw.data_ptr()), _lib.P(ev.data_ptr()),
                                                     _lib.P(cnt.data_ptr())), "getevents")
        e1.record(ctx.stream)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    ms = min(ts)
    n_ev = int(cnt.sum().item())
    reruns = int(_lib.lib().abea_ctx_getevents_reruns(ctx.handle))
    det_chunk = int(_lib.lib().abea_ctx_getevents_chunk(ctx.handle))
    # algorithmic bytes: the signal read once (4 B per sample), the event
    # records written (24 B each)
    alg = 4 * n_samples + 24 * n_ev
    res = {"workload": "getevents1024_ln5kb", "reads": b.n_reads, "samples": n_samples,
           "events": n_ev, "ms": round(ms, 3),
           "gsamples_per_s": round(n_samples / (ms * 1e-3) / 1e9, 3),
           "alg_bytes": alg, "achieved_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "chunks_rerun": reruns, "detector_chunk": det_chunk}
    log(f"[bench] getevents: {res}")
    del raw, ev, cnt
    return res, (raw_h, raw_off)


def refine_leg(args, ctx) -> dict:
    """Scaling refinement: 8,192 x 20 kb reads, scalings perturbed (shift
    +-5, scale +-10%), five graph-replayed forward + Gauss-Newton steps."""
    import numpy as np
    import torch

    from abea_amd import engine, synth
    from abea_amd._lib import PLAN_BATCHED

    b4 = synth.workload("refine8192x20kb", seed=args.seed + 4)
    true_scale, true_shift = synth.perturb_scalings(b4, seed=args.seed + 4)
    d4 = engine.DeviceBatch.upload(b4, 0)
    p4 = engine.Plan(ctx, d4, PLAN_BATCHED)
    p4.forward()
    p4.autotune(reps=1)
    p4.set_tuning(use_graph=1)
    torch.cuda.synchronize()
    s0 = d4.scalings.cpu().numpy()
    steps = 5
    t = time.perf_counter()
    ms = p4.refine(steps, 1.0, mom_start=True)
    wall = (time.perf_counter() - t) / steps
    s1 = d4.scalings.cpu().numpy()

    def err(s, col, true):
        d = np.abs(s[:, col] - true)
        return {"mean": round(float(d.mean()), 5), "median": round(float(np.median(d)), 5),
                "p90": round(float(np.quantile(d, 0.9)), 5)}

    res = {"workload": "refine8192x20kb", "steps": steps, "forwards": steps + 1,
           "start": "per read, the better scoring of the given and the MOM scalings",
           "tuning": p4.tuning(),
           "ms_per_step": round(ms, 3), "wall_ms_per_step": round(wall * 1e3, 3),
           "gcups_per_forward": round(b4.nominal_cells() * (steps + 1) / steps
                                      / (ms * 1e-3) / 1e9, 2),
           "shift_abs_err_before": err(s0, 1, true_shift),
           "shift_abs_err_after": err(s1, 1, true_shift),
           "scale_abs_err_before": err(s0, 0, true_scale),
           "scale_abs_err_after": err(s1, 0, true_scale)}
    log(f"[bench] refine: {res}")
    del p4, d4, b4
    torch.cuda.empty_cache()
    return res


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(args, b, gpu, fd_in, ge_in) -> dict:
    """The oracle (f5c semantics, test infrastructure) as the CPU baseline and
    as the checker of the GPU's headline results."""
    import numpy as np

    from abea_amd import engine, synth

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    out = {}
    rng = np.random.default_rng(args.seed)
    order = rng.permutation(b.n_reads)
    probe = b.subset(order[:32])
    t = time.perf_counter()
    oracle.align_batch(probe)
    rate = probe.nominal_cells() / (time.perf_counter() - t)
    cum = np.cumsum(b.n_bands[order] * W)
    n_s = int(min(b.n_reads, max(32, np.searchsorted(cum, rate * args.cpu_seconds))))
    longest = np.argsort(b.n_bands, kind="stable")[-64:]
    idx = np.unique(np.concatenate([order[:n_s], longest]))
    samp = b.subset(idx)
    t = time.perf_counter()
    ob = oracle.align_batch(samp, want_path=True)
    dt = time.perf_counter() - t
    cpu_gcups = samp.nominal_cells() / dt / 1e9
    threads = ob["threads"]

    # parity: every read against the GPU's headline outputs -- the timed
    # sample's results, then (untimed, unless --sample-parity) the oracle on
    # the reads the sample left out
    tally = {"identical": 0, "first_bad": None, "fields_bad": {}}

    def compare(rows, o_b):
        for j, i in enumerate(rows):
            n = int(o_b["path_len"][j])
            o = int(o_b["pair_ptr"][j])
            g = int(gpu["pair_off"][i])
            checks = {
                "path_len": int(gpu["path_len"][i]) == n,
                "n_pairs": int(gpu["n_pairs"][i]) == int(o_b["n_pairs"][j]),
                "pairs": np.array_equal(gpu["pairs"][g:g + n], o_b["path"][o:o + n]),
                "from": np.array_equal(gpu["path_from"][g:g + n], o_b["from"][o:o + n]),
                "score": gpu["score"][i] == np.float32(o_b["score"][j]),
                "avg": n == 0 or gpu["avg"][i] == o_b["avg"][j]}
            if all(checks.values()):
                tally["identical"] += 1
            else:
                for k, v in checks.items():
                    if not v:
                        tally["fields_bad"][k] = tally["fields_bad"].get(k, 0) + 1
                if tally["first_bad"] is None:
                    tally["first_bad"] = {"read": int(i),
                                          "fields": [k for k, v in checks.items() if not v]}

    compare(idx, ob)
    compared = int(idx.shape[0])
    sample_desc = f"{n_s} random reads + the 64 longest"
    if not args.sample_parity:
        rest = np.setdiff1d(np.arange(b.n_reads), idx)
        if rest.size:
            t = time.perf_counter()
            compare(rest, oracle.align_batch(b.subset(rest), want_path=True))
            compared += int(rest.size)
            sample_desc = (f"every read: the CPU timing sample ({n_s} random + the 64 longest) "
                           f"and the other {rest.size}, aligned by the oracle untimed "
                           f"({time.perf_counter() - t:.1f} s)")
    out["parity"] = {"workload": "batched16384_ln8kb", "compared": compared,
                     "identical": tally["identical"], "first_mismatch": tally["first_bad"],
                     "mismatched_fields": tally["fields_bad"],
                     "longest_compared_bases": int(b.read_len.max() if compared == b.n_reads
                                                   else b.read_len[longest].max()),
                     "sample": sample_desc,
                     "fields": ["path_len", "n_pairs", "pairs", "from", "score", "avg"]}
    log(f"[bench] parity: {out['parity']}")

    # one thread, on a smaller random sample
    n1 = int(min(b.n_reads, max(8, np.searchsorted(cum, rate / threads * args.cpu1_seconds))))
    s1 = b.subset(order[:n1])
    t = time.perf_counter()
    oracle.align_batch(s1, n_threads=1)
    dt1 = time.perf_counter() - t

    # the 64 x 2 kb reads (the reference's CPU-runnable case)
    c0 = synth.workload("cpu64x2kb", seed=args.seed + 2)
    t0s = []
    for _ in range(5):
        t = time.perf_counter()
        oracle.align_batch(c0)
        t0s.append(time.perf_counter() - t)
    t = time.perf_counter()
    oracle.align_batch(c0, n_threads=1)
    t01 = time.perf_counter() - t
    out["cpu_baseline"] = {
        "value": round(cpu_gcups, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
        "sample": f"{idx.shape[0]} reads of batched16384_ln8kb ({n_s} random + the 64 longest; "
                  f"{samp.nominal_cells() / 1e9:.2f} G cells, {dt:.1f} s)",
        "reads_per_s": round(idx.shape[0] / dt, 2),
        "cpu_model": cpu_model(), "host_threads": len(os.sched_getaffinity(0)),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "one_thread": {"gcups": round(s1.nominal_cells() / dt1 / 1e9, 4),
                       "sample": f"{n1} random reads ({s1.nominal_cells() / 1e9:.2f} G cells, "
                                 f"{dt1:.1f} s)"},
        "cpu64x2kb": {"gcups": round(c0.nominal_cells() / min(t0s) / 1e9, 4),
                      "one_thread_gcups": round(c0.nominal_cells() / t01 / 1e9, 4),
                      "reads_per_s": round(c0.n_reads / min(t0s), 1)}}
    log(f"[bench] cpu: {out['cpu_baseline']}")

    # fills: cells inside the DP matrix, from the GPU's band moves
    mv_off = engine.move_offsets(b.read_len, b.n_events)
    fills = oracle.fills_from_moves(b.n_events, b.n_kmers, gpu["moves"], mv_off)
    assert np.array_equal(fills[idx], ob["fills"]), "fills from the GPU's moves != the oracle's"
    out["fills"] = {"total": int(fills.sum()), "nominal_cells": b.nominal_cells(),
                    "fraction_of_nominal": round(float(fills.sum()) / b.nominal_cells(), 4)}

    # FP64 error of the float scores, and how often the FP64 DP's path differs
    fsub = np.unique(np.concatenate([order[:256], longest[-8:]]))
    bs = b.subset(fsub)
    o64 = oracle.align_batch(bs, mode=oracle.MODE_FP64, want_path=True)
    rel = []
    diff = 0
    for j, i in enumerate(fsub):
        s64 = o64["score"][j]
        if np.isfinite(s64):
            rel.append(abs(float(gpu["score"][i]) - s64) / abs(s64))
        n = int(o64["path_len"][j])
        g = int(gpu["pair_off"][i])
        o = int(o64["pair_ptr"][j])
        if int(gpu["path_len"][i]) != n or not np.array_equal(gpu["pairs"][g:g + n],
                                                              o64["path"][o:o + n]):
            diff += 1
    rel = np.array(rel)
    out["fp64_error"] = {"reads": int(fsub.shape[0]),
                         "score_rel_err_max": float(rel.max()),
                         "score_rel_err_median": float(np.median(rel)),
                         "path_differs_fraction": round(diff / fsub.shape[0], 4),
                         "note": "float32 GPU score vs the FP64 oracle's DP score; the FP64 DP "
                                 "places its own band"}
    log(f"[bench] fp64: {out['fp64_error']}")

    # adjoint: central differences (h = 1e-3) of the FP64 score of each read's
    # traced path, held fixed, against the GPU adjoint; and how far the FP64
    # DP's path moves when only the band is held
    sub, res, g_scal, g_model = fd_in["batch"], fd_in["res"], fd_in["g_scal"], fd_in["g_model"]
    paths = []
    for i in range(sub.n_reads):
        o, n = int(res["pair_off"][i]), int(res["path_len"][i])
        paths.append((res["pairs"][o:o + n], res["path_from"][o:o + n]))
    cnt = np.zeros(4096, np.int64)
    for i in range(sub.n_reads):
        seq = sub.read_seq(i)
        for (k, _), f in zip(*paths[i]):
            if f != 2:
                cnt[oracle.kmer_rank(seq[k:k + 6])] += 1
    top = np.argsort(cnt)[-3:]
    params = [("shift", i) for i in range(sub.n_reads)] + [("scale", i) for i in range(sub.n_reads)]
    params += [("model", int(r), c) for r in top for c in range(3)]
    h = 1e-3
    worst = 0.0
    for what in params:
        reads = [what[1]] if what[0] != "model" else range(sub.n_reads)
        f, steps = [], []
        for sign in (1, -1):
            model, scale, shift = sub.model.copy(), sub.scale.copy(), sub.shift.copy()
            arr, pos = {"shift": (shift, what[1]), "scale": (scale, what[1])}.get(
                what[0], (None, None))
            if arr is not None:
                v = np.float32(arr[pos] + sign * h)
                steps.append(float(v) - float(arr[pos]))
                arr[pos] = v
            else:
                v = np.float32(model[what[1], what[2]] + sign * h)
                steps.append(float(v) - float(model[what[1], what[2]]))
                model[what[1], what[2]] = v
            f.append(sum(oracle.path_score(paths[i][0], paths[i][1], sub.read_seq(i),
                                           sub.read_events(i), model, scale[i], shift[i])
                         for i in reads))
        fd = (f[0] - f[1]) / (steps[0] - steps[1])
        an = {"shift": lambda: g_scal[what[1], 0], "scale": lambda: g_scal[what[1], 1]}.get(
            what[0], lambda: g_model[what[1], what[2]])()
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    cells = []
    for i in range(sub.n_reads):
        rm = engine.unpack_moves(fd_in["moves"],
                                 int(engine.move_offsets(sub.read_len, sub.n_events)[i]),
                                 int(sub.n_bands[i]))
        r64 = oracle.align(sub.read_seq(i), sub.read_events(i), sub.model, float(sub.scale[i]),
                           float(sub.shift[i]), mode=oracle.MODE_FP64, replay_moves=rm)
        cells.append(len(set(map(tuple, r64["path"])) ^ set(map(tuple, paths[i][0]))))
    out["adjoint_fd"] = {"workload": "adjoint1024x10kb (8 reads)", "h": h,
                         "params": len(params), "max_rel_err": float(worst),
                         "target": 1e-3,
                         "held": "each read's traced path (FP64 path score)",
                         "fp64_dp_path_cells_differing_band_held": cells}
    log(f"[bench] adjoint fd: {out['adjoint_fd']}")

    # getevents: the oracle, one thread, on the first reads up to ~2 s of work
    raw_h, raw_off = ge_in
    t0 = time.perf_counter()
    done = 0
    for i in range(raw_off.shape[0] - 1):
        oracle.getevents(raw_h[raw_off[i]:raw_off[i + 1]])
        done = i + 1
        if time.perf_counter() - t0 > 2.0:
            break
    cpu_s = time.perf_counter() - t0
    out["getevents_cpu"] = {"gsamples_per_s": round(int(raw_off[done]) / cpu_s / 1e9, 4),
                            "cores": 1, "kind": "port",
                            "sample": f"first {done} reads ({int(raw_off[done])} samples)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true", help="headline only")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu1-seconds", type=float, default=4.0)
    ap.add_argument("--sample-parity", action="store_true",
                    help="compare only the CPU timing sample with the GPU (not every read)")
    ap.add_argument("--profile", action="store_true",
                    help="headline workload only, default tuning, no autotune; also one "
                         "1 GiB device copy to calibrate HBM byte counters")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of the captured graph")
    ap.add_argument("--leg", choices=["postalign", "getevents", "dropin", "plain", "refine",
                                      "adjoint"],
                    default=None,
                    help="run only this extra measurement (its JSON on stdout)")
    ap.add_argument("--tuning", default=None,
                    help="WPB,WPC,FUSE,OVERLAP: use this core launch shape instead of "
                         "autotuning (e.g. to profile the shape a default run picked)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: reads are independent, so N GPUs would each "
              "run this same batched plan on a length-balanced share of the reads with no "
              "exchange; the figure measured here is one GPU's. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from abea_amd import engine

    default_tune = {"core_waves_per_block": 1, "core_waves_per_cu": 24, "traceback_overlap": 128,
                    "use_graph": 1, "fuse_traceback": 1}
    if args.leg == "postalign":
        print(json.dumps(postalign_leg(args, engine.Context(0), None)), flush=True)
        return
    if args.leg == "getevents":
        print(json.dumps(getevents_leg(args, engine.Context(0))[0]), flush=True)
        return
    if args.leg == "dropin":
        print(json.dumps(dropin_leg(args)), flush=True)
        return
    if args.leg == "plain":
        print(json.dumps(plain_leg(args, engine.Context(0), default_tune)), flush=True)
        return
    if args.leg == "refine":
        print(json.dumps(refine_leg(args, engine.Context(0))), flush=True)
        return
    if args.leg == "adjoint":
        print(json.dumps(adjoint_leg(args, engine.Context(0), default_tune)[0]), flush=True)
        return

    ctx = engine.Context(0)
    out, b, gpu, tune = headline_leg(args, ctx)
    if not args.no_extras and not args.profile:
        out["plain_vs_batched"] = plain_leg(args, ctx, tune)
        out["dropins"] = dropin_leg(args)
        out["adjoint"], fd_in = adjoint_leg(args, ctx, tune)
        out["postalign"] = postalign_leg(args, ctx, tune)
        out["getevents"], ge_in = getevents_leg(args, ctx)
        out["refine"] = refine_leg(args, ctx)
        if args.cpu_seconds > 0:
            extra = cpu_leg(args, b, gpu, fd_in, ge_in)
            out.update(extra)
            out["getevents"]["cpu"] = out.pop("getevents_cpu")
            out["gpu_over_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
