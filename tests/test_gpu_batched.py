"""GPU tests of batched training (many independent problems, mixed k, one
launch per epoch) against per-problem serial CPU oracle solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(count, seed=11, scale=0.05):
    from cumf_sgd_amd import datagen

    return datagen.batched(seed=seed, count=count, users=(300, 3000), items=(100, 800), density=(0.01, 0.05))


def _oracle_group(k, prec):
    """Lanes per sample in batched launches: 16 bytes of a row per lane,
    8 to 64 lanes."""
    eb = {0: 4, 1: 2, 2: 2, 3: 1, 4: 2}[prec]
    return max(8, min(64, k * eb // 16))


@pytest.mark.parametrize("prec", [0, 1, 2, 3, 4])
def test_batched_serial_matches_oracle(gpu, prec):
    """One worker per problem: each problem's factors equal the serial oracle
    keyed by its problem index, bit for bit."""
    import cumf_sgd_amd as cs
    import oracle

    probs = _problems(6)
    seed, lam = 21, 0.05
    with cs.Batch(prec) as b:
        init = []
        for p in probs:
            u, v, r = (t.numpy() for t in (p.u, p.v, p.r))
            idx = b.add(u, v, r, p.m, p.n, p.k, lam, seed=seed)
            init.append((b.get_factors(idx, 0), b.get_factors(idx, 1)))
        b.train(cs.make_opts(schedule=cs.SERIAL, seed=seed), 0, 2)
        lr = oracle.lr_table(0.08, 0.3, 0.0, 6)
        for i, p in enumerate(probs):
            (P0, ps0), (Q0, qs0) = init[i]
            wantP0, wps = oracle.init_factors(p.m, p.k, prec, seed, 0, problem=i)
            assert np.array_equal(P0.view(np.uint8), wantP0.view(np.uint8))
            u, v, r = (t.numpy() for t in (p.u, p.v, p.r))
            Pw, Qw, psw, qsw = P0, Q0, ps0, qs0
            for ep in range(2):
                Pw, Qw, psw, qsw = oracle.sgd_serial(u, v, r, Pw, Qw, lr, ep, ep + 1, k=p.k,
                                                     G=_oracle_group(p.k, prec), prec=prec, lambda_p=lam,
                                                     lambda_q=lam, seed=seed, ps=psw, qs=qsw, problem=i)
                if prec == 3:
                    Pw, psw = oracle.fp8_rescale(Pw, psw)
                    Qw, qsw = oracle.fp8_rescale(Qw, qsw)
            P1, _ = b.get_factors(i, 0)
            Q1, _ = b.get_factors(i, 1)
            assert np.array_equal(P1.view(np.uint8), Pw.view(np.uint8)), (i, p.k)
            assert np.array_equal(Q1.view(np.uint8), Qw.view(np.uint8)), (i, p.k)


@pytest.mark.parametrize("prec,by", [(0, 1), (1, 2), (2, 1), (4, 2)])
def test_batched_order_then_serial_matches_oracle(gpu, prec, by):
    """cumf_batch_order: every problem's records stably sorted by row (1) or
    column (2) id, then one worker per problem: the factors equal the serial
    oracle run on numpy's stable argsort of the same records, bit for bit
    (consecutive samples then share rows, which the kernel keeps in
    registers instead of storing and reloading)."""
    import cumf_sgd_amd as cs
    import oracle

    probs = _problems(5, seed=13)
    seed, lam = 23, 0.05
    with cs.Batch(prec) as b:
        init = []
        for p in probs:
            u, v, r = (t.numpy() for t in (p.u, p.v, p.r))
            idx = b.add(u, v, r, p.m, p.n, p.k, lam, seed=seed)
            init.append((b.get_factors(idx, 0), b.get_factors(idx, 1)))
        b.order(by)
        b.train(cs.make_opts(schedule=cs.SERIAL, seed=seed, batch_len=32), 0, 2)
        lr = oracle.lr_table(0.08, 0.3, 0.0, 6)
        for i, p in enumerate(probs):
            (P0, ps0), (Q0, qs0) = init[i]
            u, v, r = (t.numpy() for t in (p.u, p.v, p.r))
            perm = np.argsort(u if by == 1 else v, kind="stable")
            u, v, r = u[perm], v[perm], r[perm]
            Pw, Qw = P0, Q0
            for ep in range(2):
                Pw, Qw, _, _ = oracle.sgd_serial(u, v, r, Pw, Qw, lr, ep, ep + 1, k=p.k, G=_oracle_group(p.k, prec),
                                                 prec=prec, lambda_p=lam, lambda_q=lam, seed=seed, problem=i,
                                                 batch_len=32)
            P1, _ = b.get_factors(i, 0)
            Q1, _ = b.get_factors(i, 1)
            assert np.array_equal(P1.view(np.uint8), Pw.view(np.uint8)), (i, p.k)
            assert np.array_equal(Q1.view(np.uint8), Qw.view(np.uint8)), (i, p.k)


@pytest.mark.parametrize("prec", [1, 3, 4])
def test_batched_graph_replay_equals_eager(gpu, prec):
    """use_graph=1: an epoch's batched launch (with the FP8 rescales of every
    problem) is captured once and replayed per epoch.  In serial mode the
    factors equal the eager loop's bit for bit: across two train calls (the
    device epoch counter restarts at each call's first epoch), with host
    reads between the replays, and after cumf_batch_order moves the records
    (the graph is captured again)."""
    import cumf_sgd_amd as cs

    probs = _problems(5, seed=14)
    seed = 29

    def run(graph):
        with cs.Batch(prec) as b:
            for p in probs:
                u, v, r = (t.numpy() for t in (p.u, p.v, p.r))
                b.add(u, v, r, p.m, p.n, p.k, 0.05, seed=seed)
            o = cs.make_opts(schedule=cs.SERIAL, seed=seed, batch_len=32, use_graph=graph)
            b.train(o, 0, 2)
            mid = [b.get_factors(i, 0)[0].copy() for i in range(len(probs))]
            b.train(o, 2, 3)
            b.order(cs.ORDER_COLUMNS)
            b.train(o, 3, 5)
            return mid, [b.get_factors(i, w) for i in range(len(probs)) for w in (0, 1)]

    mid_e, fin_e = run(0)
    mid_g, fin_g = run(1)
    for a, g in zip(mid_e, mid_g):
        assert np.array_equal(a.view(np.uint8), g.view(np.uint8))
    for (a, sa), (g, sg) in zip(fin_e, fin_g):
        assert np.array_equal(a.view(np.uint8), g.view(np.uint8))
        if sa is not None:
            assert np.array_equal(sa, sg)


def test_batched_hogwild_rmse_vs_serial_fp32(gpu):
    """24 problems of mixed k in BF16, 20 Hogwild epochs in one launch per
    epoch, against serial FP32 oracle solves of each problem.

    On problems this small the serial solve itself moves with the visiting
    order: over eight seeds its test RMSE spreads by up to ~0.04 (a problem
    of 2,165 ratings for 1,462 rows of 128 factors is mostly noise).  So each
    problem's GPU RMSE must lie within that spread (max over the eight serial
    seeds, plus 0.005), and the mean gap to the seed-mean over all problems
    (the systematic Hogwild + BF16 effect) must be below 0.005."""
    import cumf_sgd_amd as cs
    import oracle

    probs = _problems(24, seed=12)
    seed, epochs = 31, 20
    with cs.Batch(cs.BF16) as b:
        scales = []
        for p in probs:
            u, v, r, tu, tv, tr = (t.numpy() for t in (p.u, p.v, p.r, p.tu, p.tv, p.tr))
            sd = np.float32(np.sqrt((r.astype(np.float64) ** 2).mean() - r.astype(np.float64).mean() ** 2))
            sc = max(np.float32(1e-4), sd)
            scales.append(sc)
            rs = (r * np.float32(1.0 / sc)).astype(np.float32)
            b.add(u, v, rs, p.m, p.n, p.k, np.float32(0.05) / sc, tu=tu, tv=tv, tr=tr, seed=seed)
        b.train(cs.make_opts(schedule=cs.HOGWILD, num_workers=512, seed=seed), 0, epochs)
        lr = oracle.lr_table(0.08, 0.3, 0.0, epochs + 4)
        gaps = []
        for i, p in enumerate(probs):
            u, v, r, tu, tv, tr = (t.numpy() for t in (p.u, p.v, p.r, p.tu, p.tv, p.tr))
            sc = scales[i]
            fs = np.float32(np.sqrt(sc))
            got = b.test_rmse(i, fs, 0.0)
            rs = (r * np.float32(1.0 / sc)).astype(np.float32)
            want = []
            for s in range(seed, seed + 8):
                P, _ = oracle.init_factors(p.m, p.k, 0, s, 0, problem=i)
                Q, _ = oracle.init_factors(p.n, p.k, 0, s, 1, problem=i)
                P, Q, _, _ = oracle.sgd_serial(u, v, rs, P, Q, lr, 0, epochs, k=p.k, G=_oracle_group(p.k, 0), prec=0,
                                               lambda_p=np.float32(0.05) / sc, lambda_q=np.float32(0.05) / sc, seed=s,
                                               problem=i)
                want.append(oracle.rmse(tu, tv, tr, (P * fs).astype(np.float32), (Q * fs).astype(np.float32), 0.0))
            want = np.array(want)
            gaps.append(got - want.mean())
            print(f"problem {i}: m={p.m} n={p.n} k={p.k} nnz={p.u.numel()} gpu={got:.4f} "
                  f"oracle seeds {want.min():.4f}..{want.max():.4f}")
            assert np.isfinite(got) and got <= want.max() + 0.005
        print(f"mean gap to the serial seed-mean {np.mean(gaps):+.4f}, worst {np.max(gaps):+.4f}")
        assert np.mean(gaps) < 0.005


def test_c4_full_config_256_problems(gpu):
    """The batched configuration at full size: 256 problems (users 1K-200K,
    items 500-50K, density 0.1-2%, k in {32, 64, 128}; 677M ratings), BF16
    storage, Hogwild, one launch per epoch with the wave count and pipeline
    depth and window length the batch autotuner picks, 20 epochs.

    Checked against serial FP32 solves (batched serial mode, bit-exact with
    the oracle per problem) for the 187 problems of at most 2M ratings, and
    against FP32 Hogwild for all: every RMSE finite; no systematic gap (mean
    BF16-minus-serial below 0.005); no problem further than 0.05 from its
    serial solve (the serial solve itself moves by up to ~0.04 between seeds
    on the smallest problems).  Problems above 2M ratings against two FP32
    Hogwild seeds (which differ by up to 0.009): BF16 with stochastic
    rounding at most 0.02 above the worse seed, round-to-nearest BF16 at most
    0.03 above (it stalls where updates become small), and neither higher on
    average (mean gap below 0.005).

    Then BF16 again with every problem's records ordered by row
    (cumf_batch_order): consecutive samples of a worker share their P row,
    which the kernel keeps in registers.  The visiting order changes the
    trajectory (on most problems it ends lower than the serial solve), so the
    bars are one-sided: no problem more than 0.05 above its serial solve, no
    problem above 2M ratings more than 0.03 above the worse FP32 seed, and
    neither higher on average."""
    import time

    import torch

    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    probs = datagen.batched(seed=4, count=256, device="cuda")
    nnz = np.array([p.u.numel() for p in probs])
    assert len(probs) == 256 and nnz.sum() > 600_000_000
    scales = []
    for p in probs:
        r = p.r.double()
        sd = np.float32(np.sqrt(float((r * r).mean() - r.mean() ** 2)))
        scales.append(max(np.float32(1e-4), sd))

    def run(prec, sched, idx, seed=7, epochs=20, tune=False, order=0):
        with cs.Batch(prec) as b:
            for i in idx:
                p, sc = probs[i], scales[i]
                rs = (p.r * float(np.float32(1.0) / sc)).to(torch.float32)
                b.add(p.u, p.v, rs, p.m, p.n, p.k, float(np.float32(0.05) / sc), tu=p.tu, tv=p.tv, tr=p.tr,
                      seed=seed)
            if order:
                b.order(order)
            torch.cuda.synchronize()
            opts = cs.make_opts(schedule=sched, seed=seed)
            if tune:
                opts, ms = b.autotune(opts)
                print(f"autotuned: {opts.num_workers} waves, depth {opts.pipeline_depth}, window {opts.batch_len} "
                      f"({ms:.1f} ms/epoch)")
            t0 = time.perf_counter()
            b.train(opts, 0, epochs)
            t = time.perf_counter() - t0
            out = np.array([b.test_rmse(j, float(np.sqrt(scales[i])), 0.0) for j, i in enumerate(idx)])
        return out, nnz[idx].sum() * epochs / t

    everything = list(range(256))
    small = [i for i in everything if nnz[i] <= 2_000_000]
    bf16, rate = run(cs.BF16, cs.HOGWILD, everything, tune=True)
    fp32, rate32 = run(cs.FP32, cs.HOGWILD, everything)
    fp32b, _ = run(cs.FP32, cs.HOGWILD, everything, seed=8)
    bf16sr, rate_sr = run(cs.BF16_SR, cs.HOGWILD, everything)
    serial, _ = run(cs.FP32, cs.SERIAL, small)
    print(f"C4: BF16 Hogwild {rate / 1e9:.2f}e9 updates/s (autotuned), FP32 {rate32 / 1e9:.2f}e9, "
          f"BF16-SR {rate_sr / 1e9:.2f}e9 (defaults)")
    assert np.all(np.isfinite(bf16)) and np.all(np.isfinite(fp32)) and np.all(np.isfinite(serial))
    assert np.all(np.isfinite(bf16sr))
    gap = bf16[small] - serial
    print(f"BF16 Hogwild - FP32 serial over {len(small)} problems: mean {gap.mean():+.4f}, "
          f"max {gap.max():+.4f}, min {gap.min():+.4f}")
    assert gap.mean() < 0.005
    assert np.all(np.abs(gap) < 0.05), [(small[j], float(gap[j])) for j in np.nonzero(np.abs(gap) >= 0.05)[0]]
    large = [i for i in everything if nnz[i] > 2_000_000]
    worse32 = np.maximum(fp32[large], fp32b[large])
    d = bf16[large] - fp32[large]
    dsr = bf16sr[large] - fp32[large]
    spread = np.abs(fp32[large] - fp32b[large])
    print(f"over {len(large)} problems above 2M ratings: BF16 - FP32 max {d.max():+.4f} mean {d.mean():+.4f}; "
          f"BF16-SR - FP32 max {dsr.max():+.4f} mean {dsr.mean():+.4f}; FP32 seed spread max {spread.max():.4f}")
    # round-to-nearest BF16 drops the small late updates of the slowest
    # converging problems (its RMSE stalls up to ~0.02 above FP32 by epoch
    # 20); stochastic rounding keeps them, within twice the FP32 seed spread
    assert d.mean() < 0.005 and np.all(bf16[large] <= worse32 + 0.03)
    assert abs(dsr.mean()) < 0.005 and np.all(bf16sr[large] <= worse32 + 0.02)

    rows, rate_rows = run(cs.BF16, cs.HOGWILD, everything, order=cs.ORDER_ROWS)
    assert np.all(np.isfinite(rows))
    gr = rows[small] - serial
    dr = rows[large] - fp32[large]
    print(f"BF16 Hogwild on row-ordered records: {rate_rows / 1e9:.2f}e9 updates/s; - FP32 serial mean "
          f"{gr.mean():+.4f} max {gr.max():+.4f} min {gr.min():+.4f}; above 2M ratings - FP32 mean {dr.mean():+.4f} "
          f"max {dr.max():+.4f}")
    assert gr.mean() < 0.005 and np.all(gr < 0.05)
    assert dr.mean() < 0.005 and np.all(rows[large] <= worse32 + 0.03)
