"""GPU parity at the benchmark configurations' own sizes.

- the longest reads of the 16,384-read headline workload (log-normal median
  8 kb, clipped at 200 kb): pairs, from-codes, score and QC statistic bitwise
  against the oracle, as the reference's harness compares every read of a
  batch (GPU/align_cuda.c:315-347, GPU/dump_functions.h:151-176);
- the adjoint on 10 kb reads (the 1,024 x 10 kb workload): the autograd
  gradient against the FP64 path gradient and against central differences
  (h = 1e-3) of the FP64 best-path score with the band held fixed;
- the refinement workload's 20 kb reads under perturbed scalings: forward
  against the oracle before and after two refine steps.
"""
import concurrent.futures as cf

import numpy as np
import pytest

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


def _read_path(res, i):
    off = int(res["pair_off"][i])
    n = int(res["path_len"][i])
    return res["pairs"][off:off + n], res["path_from"][off:off + n]


def _oracle_batch_mismatches(b, res):
    """Read indices whose GPU outputs differ from the oracle's in any field."""
    import oracle
    ob = oracle.align_batch(b, want_path=True)
    bad = []
    for i in range(b.n_reads):
        n = int(ob["path_len"][i])
        off = int(ob["pair_ptr"][i])
        pairs, frm = _read_path(res, i)
        same = (int(res["path_len"][i]) == n and int(res["n_pairs"][i]) == int(ob["n_pairs"][i])
                and np.array_equal(pairs, ob["path"][off:off + n])
                and np.array_equal(frm, ob["from"][off:off + n])
                and res["score"][i] == np.float32(ob["score"][i])
                and (n == 0 or res["avg"][i] == ob["avg"][i]))
        if not same:
            bad.append(i)
    return bad


def _forward(ctx, b, form=1):
    from abea_amd import engine
    plan = engine.Plan(ctx, engine.DeviceBatch.upload(b, 0), form)
    plan.forward()
    return plan, plan.results()


@pytest.mark.parametrize("form", [1, 0], ids=["batched", "plain"])
def test_headline_tail_reads(ctx, form):
    """The 8 longest read lengths of the headline workload (up to 63.5 kb,
    ~168 k bands) and one read at the workload's 200 kb clip (~530 k bands):
    the longest walks and tracebacks of a few thousand rows per segment, in
    the batched form and in the reference's shape (one launch pair per
    read)."""
    from abea_amd import synth
    L = np.sort(synth.workload_lengths("batched16384_ln8kb", seed=0))[-8:]
    assert L[-1] > 60000
    L = np.concatenate([L, [200000]])
    b = synth.generate(L, seed=101)
    print(f"[tail] form {form}: lengths {L.tolist()}, events {b.n_events.tolist()}", flush=True)
    plan, res = _forward(ctx, b, form)
    assert _oracle_batch_mismatches(b, res) == []
    plan.close()


def test_plain_form_on_plain_workload_reads(ctx):
    """96 reads drawn from the 4,096-read 5 kb workload (the one the plain
    form is timed on) in the plain form, and the same reads in the batched
    form: every field of every read equal to the oracle's."""
    from abea_amd import synth
    lens = synth.workload_lengths("plain4096_ln5kb", seed=1)
    rng = np.random.default_rng(23)
    b = synth.generate(rng.choice(lens, 96, replace=False), seed=23)
    for form in (0, 1):
        print(f"[plain] form {form}: {b.n_reads} reads, {int(b.read_len.sum())} bases", flush=True)
        plan, res = _forward(ctx, b, form)
        assert _oracle_batch_mismatches(b, res) == [], form
        plan.close()


def _fd_params(b, res):
    """(shift/scale of each read) + three k-mers x (mu, sigma, lambda): the
    k-mers with the most D/U cells over the batch's paths."""
    import oracle
    cnt = np.zeros(4096, np.int64)
    for i in range(b.n_reads):
        pairs, frm = _read_path(res, i)
        seq = b.read_seq(i)
        for (k, _), f in zip(pairs, frm):
            if f != 2:
                cnt[oracle.kmer_rank(seq[k:k + 6])] += 1
    top = np.argsort(cnt)[-3:]
    ps = [("shift", i) for i in range(b.n_reads)] + [("scale", i) for i in range(b.n_reads)]
    ps += [("model", int(r), c) for r in top for c in range(3)]
    return ps


def _perturb(b, what, sign, h):
    """(model, scale[], shift[], step) with one float32 input moved by
    sign * h (the inputs are float32, so the step actually taken is
    returned)."""
    model, scale, shift = b.model.copy(), b.scale.copy(), b.shift.copy()
    if what[0] == "shift":
        v = np.float32(shift[what[1]] + sign * h)
        step = float(v) - float(shift[what[1]])
        shift[what[1]] = v
    elif what[0] == "scale":
        v = np.float32(scale[what[1]] + sign * h)
        step = float(v) - float(scale[what[1]])
        scale[what[1]] = v
    else:
        r, c = what[1], what[2]
        v = np.float32(model[r, c] + sign * h)
        step = float(v) - float(model[r, c])
        model[r, c] = v
    return model, scale, shift, step


def test_adjoint_10kb_reads_vs_fd(ctx):
    """Eight 10 kb reads through torch.autograd (abea_score).

    1. The backward equals the FP64 path gradient (partial derivatives) along
       the GPU's traced paths.
    2. Central differences at h = 1e-3 of the FP64 score of each read's
       traced path, held fixed, agree with the backward within 1e-3 relative
       for every read's shift and scale and three k-mers' level_mean,
       level_stdv and level_log_stdv (the path score is quadratic in shift,
       scale and level_mean, so these differences are exact but for
       rounding).
    3. With only the band held (the GPU's moves replayed) the FP64 DP's best
       path is not the float path: float rounding moves near-ties, so on
       these reads the two paths differ in a few cells, and a step of 1e-3 in
       scale moves every cell's emission by ~0.1 and switches the DP's path.
       Where the FP64 DP path stays put within +-h its differences equal that
       path's own gradient (asserted); the gap to the float path's gradient
       is reported.
    4. The GPU's own float32 score differenced at h = 1e-3 with the band
       replayed is reported, not asserted: each forward carries ~0.1-1 of
       float rounding on scores of ~7e4, which h = 1e-3 turns into errors
       far above 1e-3."""
    import oracle
    from abea_amd import engine, synth
    from abea_amd.autograd import abea_score
    b = synth.workload("adjoint1024x10kb", seed=3, n=8)
    plan = engine.Plan(ctx, engine.DeviceBatch.upload(b, 0), 1)
    scal = torch.tensor(b.scalings(), device="cuda", requires_grad=True)
    model = torch.tensor(b.model.astype(np.float32), device="cuda", requires_grad=True)
    w = torch.linspace(0.75, 1.25, b.n_reads, device="cuda")
    score = abea_score(plan, scal, model)
    (score * w).sum().backward()
    res = plan.results()
    moves = plan.moves()
    mv_off = engine.move_offsets(b.read_len, b.n_events)
    wn = w.cpu().numpy().astype(np.float64)
    gs = scal.grad.cpu().numpy().astype(np.float64)
    gm = model.grad.cpu().numpy().astype(np.float64)
    paths = [_read_path(res, i) for i in range(b.n_reads)]
    # 1. the backward against the FP64 path gradient along the GPU's paths
    ref = np.zeros((4096, 3))
    for i in range(b.n_reads):
        gr, g_mu, g_sg, g_lam = oracle.path_gradient(paths[i][0], paths[i][1], b.read_seq(i),
                                                     b.read_events(i), b.model, b.scale[i],
                                                     b.shift[i], weight=wn[i], partials=True)
        np.testing.assert_allclose(gs[i, 0], gr[1], rtol=1e-5)
        np.testing.assert_allclose(gs[i, 1], gr[0], rtol=1e-5, atol=1e-3)
        ref[:, 0] += g_mu
        ref[:, 1] += g_sg
        ref[:, 2] += g_lam
    np.testing.assert_allclose(gm, ref, rtol=1e-6, atol=1e-5)

    def analytic(what):
        if what[0] == "shift":
            return gs[what[1], 1]
        if what[0] == "scale":
            return gs[what[1], 0]
        return gm[what[1], what[2]]

    def reads_of(what):
        return [what[1]] if what[0] in ("shift", "scale") else range(b.n_reads)

    h = 1e-3
    params = _fd_params(b, res)

    # 2. FD of the FP64 score of the held GPU path
    def fd_path(what):
        f, steps = [], []
        for sign in (1, -1):
            m_, sc_, sh_, st = _perturb(b, what, sign, h)
            f.append(sum(wn[i] * oracle.path_score(paths[i][0], paths[i][1], b.read_seq(i),
                                                   b.read_events(i), m_, sc_[i], sh_[i])
                         for i in reads_of(what)))
            steps.append(st)
        return (f[0] - f[1]) / (steps[0] - steps[1])

    worst = 0.0
    for what in params:
        an = analytic(what)
        rel = abs(fd_path(what) - an) / max(abs(an), 1e-12)
        worst = max(worst, rel)
        assert rel <= 1e-3, (what, an, rel)
    print(f"[adjoint] FD of the held path's FP64 score: {len(params)} params, "
          f"max rel err {worst:.2e}", flush=True)

    # 3. FD of the FP64 DP score with the band held at the GPU's moves
    rm = [engine.unpack_moves(moves, int(mv_off[i]), int(b.n_bands[i]))
          for i in range(b.n_reads)]

    def dp64(i, model_, scale_, shift_):
        return oracle.align(b.read_seq(i), b.read_events(i), model_, float(scale_[i]),
                            float(shift_[i]), mode=oracle.MODE_FP64, replay_moves=rm[i])

    with cf.ThreadPoolExecutor(8) as ex:
        base = list(ex.map(lambda i: dp64(i, b.model, b.scale, b.shift), range(b.n_reads)))
    n_diff = [len(set(map(tuple, base[i]["path"])) ^ set(map(tuple, paths[i][0])))
              for i in range(b.n_reads)]

    def fd_dp(what):
        f, steps = [], []
        for sign in (1, -1):
            m_, sc_, sh_, st = _perturb(b, what, sign, h)
            tot = 0.0
            for i in reads_of(what):
                r = dp64(i, m_, sc_, sh_)
                if not np.array_equal(r["path"], base[i]["path"]):
                    return None  # the DP's best path switches inside [-h, h]
                tot += wn[i] * r["score"]
            f.append(tot)
            steps.append(st)
        return (f[0] - f[1]) / (steps[0] - steps[1])

    with cf.ThreadPoolExecutor(8) as ex:
        fds = list(ex.map(fd_dp, params))
    stable, gap = 0, 0.0
    for what, v in zip(params, fds):
        if v is None:
            continue
        stable += 1
        g64 = 0.0
        for i in reads_of(what):
            gr, g_mu, g_sg, g_lam = oracle.path_gradient(base[i]["path"], base[i]["from"],
                                                         b.read_seq(i), b.read_events(i), b.model,
                                                         b.scale[i], b.shift[i], weight=wn[i],
                                                         partials=True)
            if what[0] == "shift":
                g64 += gr[0]
            elif what[0] == "scale":
                g64 += gr[1]
            else:
                g64 += (g_mu, g_sg, g_lam)[what[2]][what[1]]
        assert abs(v - g64) <= 1e-6 * abs(g64) + 1e-6, (what, v, g64)
        an = analytic(what)
        gap = max(gap, abs(v - an) / max(abs(an), 1e-12))
    print(f"[adjoint] band held: FP64 DP path differs from the float path in {n_diff} cells; "
          f"DP path stable within +-h for {stable}/{len(params)} params; max rel gap of their "
          f"FD to the float path's gradient {gap:.2e}", flush=True)

    # 4. the GPU's float32 score differenced with the band replayed (reported)
    plan.set_replay(moves)
    gpu_worst = 0.0
    for what in [("shift", 0), ("scale", 0)]:
        f, steps = [], []
        for sign in (1, -1):
            _, sc_, sh_, st = _perturb(b, what, sign, h)
            sp = np.stack([sc_, sh_, b.var, b.log_var], 1).astype(np.float32)
            plan.set_params(torch.tensor(sp, device="cuda"))
            plan.forward()
            f.append(float(plan.results()["score"][what[1]]))
            steps.append(st)
        v = (f[0] - f[1]) / (steps[0] - steps[1])
        an = analytic(what) / wn[what[1]]
        gpu_worst = max(gpu_worst, abs(v - an) / abs(an))
    print(f"[adjoint] GPU float32 FD at h=1e-3 (band replayed): max rel err {gpu_worst:.2e}",
          flush=True)
    plan.close()


def test_refine_20kb_reads_vs_oracle(ctx):
    """Eight 20 kb reads of the refinement workload with scalings perturbed
    (shift +-5, scale +-10%): the forward equals the oracle before and after
    two refine steps (the walks spend long stretches in the trim column and
    below the last event under such scalings)."""
    from abea_amd import engine, synth
    b = synth.workload("refine8192x20kb", seed=4, n=8)
    true_scale, true_shift = synth.perturb_scalings(b, seed=4)
    plan, res = _forward(ctx, b)
    assert _oracle_batch_mismatches(b, res) == []
    plan.refine(2, 1.0)
    sc = plan.batch.scalings.cpu().numpy()
    plan.forward()
    res2 = plan.results()
    b2 = b.subset(np.arange(b.n_reads))
    b2.scale = sc[:, 0].copy()
    b2.shift = sc[:, 1].copy()
    assert not np.array_equal(b2.scale, b.scale)
    assert _oracle_batch_mismatches(b2, res2) == []
    print(f"[refine] |shift error| {np.abs(b.shift - true_shift).round(2).tolist()} -> "
          f"{np.abs(b2.shift - true_shift).round(2).tolist()}", flush=True)
    assert np.all(np.abs(b2.shift - true_shift) < 1.5), (b2.shift, true_shift)
    plan.close()
