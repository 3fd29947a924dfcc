"""ABEA benchmark on one MI355X.

Headline: batched persistent ABEA (length-sorted queue, persistent core that
stages, walks and traces back each read, replayed as a hipGraph, autotuned
launch shapes) over the
16,384-read long-tailed workload (log-normal median 8 kb, clipped 500 b-200 kb,
~150 Mbases, ~40 G nominal cells), inputs resident in HBM as f5c event_t
records.  A "step" is one full forward over that batch.

Also reported (same JSON line): the plain per-read-launch form and the batched
form on a sample of the 4,096-read log-normal (median 5 kb) workload; the
adjoint/forward time ratio on 1,024 x 10 kb reads; the scaling-refinement loop
on 8,192 x 20 kb reads (five graph-replayed forward + Gauss-Newton steps from
perturbed scalings, with the recovered-scaling error); the CPU baseline (the
oracle, OpenMP over reads on this host) on a bounded sample.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "abea-opencl_amd"))

HBM_PEAK_GBS = 8000.0
W = 100
# HBM bytes per launch of the core kernel on the headline workload, from
# rocprofv3 FETCH_SIZE + WRITE_SIZE passes scaled by a 1 GiB copy's readings
# (profiles/abea_kernels_batched16384_ln8kb.md gives the runs), for the
# fused-traceback launch shape the autotuner picks here.
TRAFFIC_BYTES = {"abea_core_kernel<1, true>": 49442908774}
# VALU wave-instructions per launch of the same kernel on the same workload
# (rocprofv3 SQ_INSTS_VALU, same file).  The core is bound by VALU issue: a
# SIMD issues one wave64 VALU instruction per 4 cycles.
VALU_INSTS = {"abea_core_kernel<1, true>": 2.086e10}
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

    def best_ms(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return 1e3 * min(ts)

    fwd_ms = best_ms(p5.forward)
    mom_ms = best_ms(lambda: p5.estimate_scalings())
    full_ms = best_ms(lambda: p5.postalign())
    nrec_ms = best_ms(lambda: p5.postalign(records=False))
    pr = engine.postalign_results(p5, p5.postalign())
    n_pairs = int(p5.n_pairs[:b5.n_reads].sum().item())
    n_rec = int(pr["n_alignment"].sum())
    n_km = int(np.maximum(b5.read_len.astype(np.int64) - 5, 0).sum())
    # algorithmic bytes: pairs read (8 B each), records written (32 B),
    # map entries written (8 B per k-mer), an event mean read (4 B) per
    # record (an upper bound: only 'M' records need one); model rows are
    # cache-resident
    post_bytes = 8 * n_pairs + 32 * n_rec + 8 * n_km + 4 * n_rec
    flags = np.bincount(pr["read_stat_flag"], minlength=3)
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
                   "failed_alignment": int(flags[2])}}
    log(f"[bench] postalign: {res}")
    del p5, d5, b5, pr
    return res


def getevents_leg(args, ctx) -> dict:
    """f5c's step before MOM and align(): event detection on the raw signals
    (pA) of 1,024 reads of the 5 kb log-normal workload, resident in HBM.
    The oracle's serial time on a bounded sample of the same reads is the CPU
    baseline."""
    import numpy as np
    import torch

    from abea_amd import _lib, engine, synth

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline only

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
    # algorithmic bytes: the signal read once (4 B per sample), the event
    # records written (24 B each)
    alg = 4 * n_samples + 24 * n_ev
    # CPU: the oracle, one thread, on the first reads up to ~2 s of work
    t0 = time.perf_counter()
    done = 0
    for i in range(b.n_reads):
        oracle.getevents(raw_h[raw_off[i]:raw_off[i + 1]])
        done = i + 1
        if time.perf_counter() - t0 > 2.0:
            break
    cpu_s = time.perf_counter() - t0
    cpu_samples = int(raw_off[done])
    res = {"workload": "getevents1024_ln5kb", "reads": b.n_reads, "samples": n_samples,
           "events": n_ev, "ms": round(ms, 3),
           "gsamples_per_s": round(n_samples / (ms * 1e-3) / 1e9, 3),
           "alg_bytes": alg, "achieved_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "chunks_rerun": reruns,
           "cpu": {"gsamples_per_s": round(cpu_samples / cpu_s / 1e9, 4), "cores": 1,
                   "kind": "port", "sample": f"first {done} reads ({cpu_samples} samples)"}}
    log(f"[bench] getevents: {res}")
    del raw, ev, cnt
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true", help="headline only")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--plain-reads", type=int, default=256)
    ap.add_argument("--profile", action="store_true",
                    help="headline workload only, default tuning, no autotune; also one "
                         "1 GiB device copy to calibrate HBM byte counters")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of the captured graph")
    ap.add_argument("--leg", choices=["postalign", "getevents"], default=None,
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

    import numpy as np
    import torch

    from abea_amd import engine, synth
    from abea_amd._lib import PLAN_BATCHED, PLAN_PLAIN

    if args.leg == "postalign":
        print(json.dumps(postalign_leg(args, engine.Context(0), None)), flush=True)
        return
    if args.leg == "getevents":
        print(json.dumps(getevents_leg(args, engine.Context(0))), flush=True)
        return

    t = time.perf_counter()
    wl = "batched16384_ln8kb"
    b = synth.workload(wl, seed=args.seed, with_aos=True)
    cells = b.nominal_cells()
    log(f"[bench] generated {wl}: {b.n_reads} reads, {int(b.read_len.sum()) / 1e6:.1f} Mb, "
        f"{int(b.n_events.sum()) / 1e6:.1f} M events, {cells / 1e9:.2f} G cells "
        f"in {time.perf_counter() - t:.1f}s")
    ctx = engine.Context(0)
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
        plan.autotune(reps=1)
        log(f"[bench] autotuned in {time.perf_counter() - t:.1f}s")
    tune = plan.tuning()
    log(f"[bench] tuning {tune}")

    # headline: graph replay, whole steps
    plan.set_tuning(use_graph=0 if args.eager else 1)
    for _ in range(args.warmup):
        plan.forward()
    sec = timed_steps(plan, args.steps, torch)
    gcups = cells / sec / 1e9
    log(f"[bench] batched {'eager' if args.eager else 'graph'}: {sec * 1e3:.2f} ms/step, "
        f"{gcups:.1f} GCUPS")

    # per-kernel device times over the same number of eager steps
    plan.set_tuning(use_graph=0)
    plan.forward()
    kt = np.zeros(3)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        plan.forward()
        kt += np.array(plan.times()[:3])
    kt /= args.steps
    log(f"[bench] kernel ms queue_reset/core/post: {kt.round(3).tolist()}")
    res = plan.results()
    qc_pass = int((res["n_pairs"] > 0).sum())
    walk_b, tb_b = alg_bytes(b, res["path_len"])
    # the dominant kernel is the core; fused, it also does the traceback
    core_bytes = walk_b + (tb_b if tune["fuse_traceback"] else 0)
    achieved = core_bytes / (kt[1] * 1e-3) / 1e9
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
        "kernel_ms": {"queue_reset": round(kt[0], 3), "core": round(kt[1], 3),
                      "post": round(kt[2], 3)},
        "qc_pass_reads": qc_pass,
        "roofline": {"bound": "hbm", "kernel": kernel_name(tune), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_BYTES.get(kernel_name(tune)),
                     "alg_bytes_per_launch": core_bytes},
    }
    vi = VALU_INSTS.get(kernel_name(tune))
    if vi is not None:
        # VALU issue rate against the ceiling of one wave instruction per SIMD
        # per 4 cycles at the peak clock
        rate = vi / (kt[1] * 1e-3)
        peak = N_SIMD * CLOCK_HZ / 4
        out["issue"] = {"bound": "valu", "valu_insts_per_launch": vi,
                        "valu_insts_per_band": round(vi / int(b.n_bands.sum()), 2),
                        "achieved": round(rate / 1e9, 1), "peak": round(peak / 1e9, 1),
                        "unit": "G wave-instructions/s", "frac": round(rate / peak, 4)}
    del plan, db
    torch.cuda.empty_cache()

    if not args.no_extras and not args.profile:
        # plain per-read launches vs batched on the same sample of the
        # 4,096-read log-normal (median 5 kb) workload
        L1 = synth.workload_lengths("plain4096_ln5kb", seed=args.seed)
        b1 = synth.generate(L1, seed=args.seed + 1)
        sub = b1.subset(np.arange(min(args.plain_reads, b1.n_reads)))
        dsub = engine.DeviceBatch.upload(sub, 0)
        p_plain = engine.Plan(ctx, dsub, PLAN_PLAIN)
        p_plain.forward()
        s_plain = timed_steps(p_plain, 1, torch)
        p_b = engine.Plan(ctx, dsub, PLAN_BATCHED)
        p_b.forward()
        s_b = timed_steps(p_b, 3, torch)
        d1 = engine.DeviceBatch.upload(b1, 0)
        p_full = engine.Plan(ctx, d1, PLAN_BATCHED)
        p_full.set_tuning(**{k: v for k, v in tune.items()})
        p_full.forward()
        s_full = timed_steps(p_full, 3, torch)
        out["plain_vs_batched"] = {
            "workload": "plain4096_ln5kb", "sample_reads": sub.n_reads,
            "plain_gcups": round(sub.nominal_cells() / s_plain / 1e9, 3),
            "batched_gcups_same_sample": round(sub.nominal_cells() / s_b / 1e9, 3),
            "batched_gcups_all_4096": round(b1.nominal_cells() / s_full / 1e9, 3),
            "plain_reads_per_s": round(sub.n_reads / s_plain, 2),
            "batched_reads_per_s_all_4096": round(b1.n_reads / s_full, 1),
        }
        log(f"[bench] plain/batched: {out['plain_vs_batched']}")
        del p_plain, p_b, p_full, d1, dsub

        # adjoint / forward on 1,024 x 10 kb
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
        for _ in range(3):
            p3.adjoint(gsc)
        torch.cuda.synchronize()
        s_adj = (time.perf_counter() - t) / 3
        out["adjoint"] = {"workload": "adjoint1024x10kb", "forward_ms": round(s_fwd * 1e3, 3),
                          "adjoint_ms": round(s_adj * 1e3, 3),
                          "adjoint_over_forward": round(s_adj / s_fwd, 4)}
        log(f"[bench] adjoint: {out['adjoint']}")
        del p3, d3

        out["postalign"] = postalign_leg(args, ctx, tune)
        out["getevents"] = getevents_leg(args, ctx)
        torch.cuda.empty_cache()

        # scaling refinement: 8,192 x 20 kb reads, scalings perturbed (shift
        # +-5, scale +-10%), five graph-replayed forward + Gauss-Newton steps
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
        ms = p4.refine(steps, 1.0)
        wall = (time.perf_counter() - t) / steps
        s1 = d4.scalings.cpu().numpy()
        out["refine"] = {
            "workload": "refine8192x20kb", "steps": steps, "tuning": p4.tuning(),
            "ms_per_step": round(ms, 3),
            "wall_ms_per_step": round(wall * 1e3, 3),
            "gcups_per_step": round(b4.nominal_cells() / (ms * 1e-3) / 1e9, 2),
            "shift_abs_err_before": {"mean": round(float(np.abs(s0[:, 1] - true_shift).mean()), 4),
                                     "median": round(float(np.median(np.abs(s0[:, 1] - true_shift))), 4)},
            "shift_abs_err_after": {"mean": round(float(np.abs(s1[:, 1] - true_shift).mean()), 4),
                                    "median": round(float(np.median(np.abs(s1[:, 1] - true_shift))), 4)},
            "scale_abs_err_before": {"mean": round(float(np.abs(s0[:, 0] - true_scale).mean()), 5),
                                     "median": round(float(np.median(np.abs(s0[:, 0] - true_scale))), 5)},
            "scale_abs_err_after": {"mean": round(float(np.abs(s1[:, 0] - true_scale).mean()), 5),
                                    "median": round(float(np.median(np.abs(s1[:, 0] - true_scale))), 5)}}
        log(f"[bench] refine: {out['refine']}")
        del p4, d4, b4
        torch.cuda.empty_cache()

        # CPU baseline: the oracle (f5c semantics), OpenMP over reads, on a
        # bounded random sample of the headline workload
        if args.cpu_seconds > 0:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            rng = np.random.default_rng(args.seed)
            order = rng.permutation(b.n_reads)
            probe = b.subset(order[:32])
            t = time.perf_counter()
            _, _, used, _ = oracle.align_batch(probe)
            rate = probe.nominal_cells() / (time.perf_counter() - t)
            want = rate * args.cpu_seconds
            cum = np.cumsum(b.n_bands[order] * W)
            n_s = int(min(b.n_reads, max(32, np.searchsorted(cum, want))))
            samp = b.subset(order[:n_s])
            t = time.perf_counter()
            _, _, used, _ = oracle.align_batch(samp)
            dt = time.perf_counter() - t
            cpu_gcups = samp.nominal_cells() / dt / 1e9
            out["cpu_baseline"] = {
                "value": round(cpu_gcups, 4), "unit": "GCUPS", "cores": used, "kind": "port",
                "sample": f"{n_s} random reads of batched16384_ln8kb "
                          f"({samp.nominal_cells() / 1e9:.2f} G cells, {dt:.1f} s)",
                "reads_per_s": round(n_s / dt, 2)}
            out["gpu_over_cpu"] = round(gcups / cpu_gcups, 1)
            log(f"[bench] cpu: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
