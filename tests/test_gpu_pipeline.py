"""GPU tests of the reference-seam entry points: the CLI, cumf_sgd_train
(sgd_train) and cumf_sgd_update_k128 (sgd_update_k128), with the model file
scored by the reference's compiled mf-predict."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(seed=5, m=3000, n=800, nnz=200_000, n_test=20_000):
    from cumf_sgd_amd import datagen

    R = datagen.uniform(seed, m, n, nnz, n_test)
    return R


def _nodes(u, v, r):
    import cumf_sgd_amd as cs

    x = np.zeros(len(u), cs.NODE_DTYPE)
    x["u"], x["v"], x["rate"] = u, v, r
    return x


@pytest.mark.parametrize("grid", [(1, 1, 1, 1), (2, 2, 1, 1), (2, 2, 2, 2)])
def test_sgd_train_pipeline_and_model_file(gpu, tmp_path, grid):
    import cumf_sgd_amd as cs
    import oracle

    R = _data()
    u, v, r, tu, tv, tr = R.numpy()
    m, n = int(u.max()) + 1, int(v.max()) + 1
    p = cs.make_param(k=64, num_workers=256, u_grid=grid[0], v_grid=grid[1], x_grid=grid[2], y_grid=grid[3],
                      num_iters=8)
    P, Q, b, rmse = cs.sgd_train(p, _nodes(u, v, r), m, n, precision=cs.FP16, seed=3, test=_nodes(tu, tv, tr))
    assert np.all(np.isfinite(P)) and np.all(np.isfinite(Q))
    assert rmse[-1] < 0.95 * rmse[0]
    # b = ave/std of the training ratings (sgd.cu:679)
    ave, std = oracle.collect_data(_nodes(u, v, r))
    assert b == np.float32(np.float32(ave) / np.float32(std))
    # the exported model scores the same through the oracle and the reference's evaluator
    want = oracle.rmse(tu, tv, tr, P, Q, np.float32(b))
    assert abs(want - rmse[-1]) <= 1e-9
    model = str(tmp_path / "model")
    cs.save_model(model, P, Q, b)
    if oracle.have_mf_predict():
        from cumf_sgd_amd import datagen

        test = str(tmp_path / "test.bin")
        datagen.write_triplets(test, tu, tv, tr)
        assert oracle.mf_predict_rmse(test, model) == f"{rmse[-1]:.4f}"


@pytest.mark.parametrize("rounds", [0, 4])
def test_sgd_train_stratum_schedule(gpu, rounds):
    """cumf_sgd_train on a 1x1 grid with the stratum schedule (24 x 24
    blocks, one or four passes over the diagonals per epoch) and with the
    library's default for this problem, whose 0.5 MB of factors fit in the
    L2s: the strata with four passes per epoch.  All train and stay finite,
    and the default ends within 0.005 of the explicit strata."""
    import cumf_sgd_amd as cs

    R = _data()
    u, v, r, tu, tv, tr = R.numpy()
    m, n = int(u.max()) + 1, int(v.max()) + 1
    p = cs.make_param(k=64, num_workers=256, num_iters=8)
    with cs.Context(0) as c:
        c.set_ratings(u, v, r, m, n)
        assert c.default_schedule(64, cs.FP16) == (cs.STRATUM, 4, 0)
    _, _, _, dflt = cs.sgd_train(p, _nodes(u, v, r), m, n, precision=cs.FP16, seed=3, test=_nodes(tu, tv, tr))
    o = cs.make_opts(k=64, schedule=cs.STRATUM, stratum_rounds=rounds)
    P, Q, b, st = cs.sgd_train(p, _nodes(u, v, r), m, n, precision=cs.FP16, seed=3, test=_nodes(tu, tv, tr),
                               opts=o)
    print(f"rounds={rounds} stratum {st[-1]:.4f} default (strata, 4 passes) {dflt[-1]:.4f}")
    assert np.all(np.isfinite(P)) and np.all(np.isfinite(Q))
    assert st[-1] < 0.95 * st[0] and dflt[-1] < 0.95 * dflt[0]
    assert dflt[-1] <= st[-1] + 0.005


@pytest.mark.parametrize("g", [1, 2])
def test_sgd_update_k128_seam(gpu, g):
    """The reference's host pipeline (restated by the oracle: maps, grid,
    scale with id localisation, FP16 init) feeding the drop-in for
    sgd_update_k128, on a 1x1 grid (stratum schedule inside) and a 2x2 grid
    (the reference's block order); outputs are consistent and the model
    trains."""
    import cumf_sgd_amd as cs
    import oracle

    R = _data(nnz=100_000, n_test=10_000)
    u, v, r, tu, tv, tr = R.numpy()
    m, n = int(u.max()) + 1, int(v.max()) + 1
    nodes = _nodes(u, v, r)
    ave, std = oracle.collect_data(nodes)
    scale = max(np.float32(1e-4), np.float32(std))
    pm, qm = oracle.gen_random_map(7, m), oracle.gen_random_map(7, n)
    shuf = oracle.shuffle_problem(nodes, pm, qm)
    ux = vy = g
    grid, gs, off = oracle.grid_problem(shuf, m, n, ux, vy)
    us, vs = oracle.grid_segments(m, n, ux, vy)
    local = oracle.scale_problem(grid, np.float32(1.0 / scale), ux, vy, us, vs)
    blocks = [local[off[b]:off[b + 1]] for b in range(ux * vy)]
    k = 128
    hp, _ = oracle.init_factors(ux * us, k, oracle.F16, 7, 0)
    hq, _ = oracle.init_factors(vy * vs, k, oracle.F16, 7, 1)
    p = cs.make_param(k=k, num_workers=128, u_grid=g, v_grid=g, num_iters=5, lambda_p=0.05 / scale)
    hp2, hq2 = hp.copy(), hq.copy()
    fp, fq = cs.sgd_update_k128(p, m, n, us, vs, blocks, hp2, hq2, seed=7)
    assert not np.array_equal(hp2, hp)
    assert np.array_equal(fp, oracle.widen(hp2, None, oracle.F16))
    assert np.array_equal(fq, oracle.widen(hq2, None, oracle.F16))
    fs = np.float32(np.sqrt(scale))
    before = oracle.rmse(pm[tu], qm[tv], tr, oracle.widen(hp, None, 1)[:m] * fs, oracle.widen(hq, None, 1)[:n] * fs, 0)
    after = oracle.rmse(pm[tu], qm[tv], tr, fp[:m] * fs, fq[:n] * fs, 0)
    assert after < before * 0.5


def test_cli_end_to_end(gpu, tmp_path):
    import cumf_sgd_amd as cs
    import oracle
    from cumf_sgd_amd import datagen

    R = _data(nnz=150_000, n_test=15_000)
    u, v, r, tu, tv, tr = R.numpy()
    tr_path, te_path = str(tmp_path / "train.bin"), str(tmp_path / "test.bin")
    datagen.write_triplets(tr_path, u, v, r)
    datagen.write_triplets(te_path, tu, tv, tr)
    model = str(tmp_path / "model")
    out = subprocess.run([cs.CLI_PATH, "-g", "0", "-l", "0.05", "-a", "0.08", "-b", "0.3", "-s", "200", "-k", "128",
                          "-t", "6", "--seed", "4", "-p", te_path, tr_path, model], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [l for l in out.stdout.splitlines() if "test RMSE" in l]
    assert len(lines) == 6
    last = lines[-1].split()[-1]
    size = os.path.getsize(model)
    assert size == 20 + 4 * (int(u.max()) + 1 + int(v.max()) + 1) * 128
    if oracle.have_mf_predict():
        assert oracle.mf_predict_rmse(te_path, model) == last


def test_cli_serial_mode_is_deterministic(gpu, tmp_path):
    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    R = _data(nnz=30_000, n_test=0)
    u, v, r, *_ = R.numpy()
    tr_path = str(tmp_path / "train.bin")
    datagen.write_triplets(tr_path, u, v, r)
    outs = []
    for i in range(2):
        model = str(tmp_path / f"m{i}")
        res = subprocess.run([cs.CLI_PATH, "-k", "64", "-t", "2", "--serial", "--seed", "9", tr_path, model],
                             capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        outs.append(open(model, "rb").read())
    assert outs[0] == outs[1]


@pytest.mark.parametrize("k", [80, 128])
def test_predict_cli_matches_reference_mf_predict(gpu, tmp_path, k):
    """cumf_predict (GPU) prints what the reference's compiled mf-predict
    prints for every supported -e, on model files of the reference's default
    k=80 and k=128, with out-of-range test ids; set_model + test_eval agree
    with the oracle."""
    import cumf_sgd_amd as cs
    import oracle
    from cumf_sgd_amd import datagen

    rng = np.random.default_rng(k)
    m, n = 700, 300
    P = np.abs(rng.normal(0.3, 0.15, (m, k))).astype(np.float32)
    Q = np.abs(rng.normal(0.3, 0.15, (n, k))).astype(np.float32)
    b = np.float32(3.4)
    tu = rng.integers(-1, m + 2, 20_000).astype(np.int32)
    tv = rng.integers(0, n, 20_000).astype(np.int32)
    tr = rng.integers(1, 6, 20_000).astype(np.float32)
    model, test = str(tmp_path / "model"), str(tmp_path / "test.bin")
    cs.save_model(model, P, Q, b)
    datagen.write_triplets(test, tu, tv, tr)
    label = {0: "RMSE", 1: "MAE", 2: "GKL", 5: "LOGLOSS", 6: "ACCURACY"}
    with cs.Context(0) as c:
        c.set_model(P, Q)
        c.set_test(tu, tv, tr)
        for crit in (0, 1, 2, 5, 6):
            got = c.test_eval(crit, 1.0, b)
            want = oracle.evaluate(crit, tu, tv, tr, P, Q, b)
            assert abs(got - want) <= (1e-6 if crit in (2, 5) else 1e-12) * abs(want), (crit, got, want)
    for crit in (0, 1, 2, 5, 6):
        out = subprocess.run([cs.PREDICT_PATH, "-e", str(crit), test, model], capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        line = out.stdout.strip().splitlines()[-1]
        assert line.startswith(label[crit] + " = ")
        value = line.split("=", 1)[1].strip()
        want = f"{oracle.evaluate(crit, tu, tv, tr, P, Q, b):.4f}"
        assert value == want, (crit, value, want)
        if oracle.have_mf_predict():
            assert value == oracle.mf_predict_eval(test, model, crit)


def test_streaming_reader_matches_read_problem(gpu, tmp_path):
    """cumf_read_ratings / cumf_read_test: the records of a single file (with
    a trailing partial record) and of shards path0..path3 land in the device
    arrays in file order, bit for bit, with read_problem's m and n; negative
    ids and missing files are errors."""
    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    rng = np.random.default_rng(21)
    # single file spanning several reader chunks (4M records each)
    nnz = 9_000_001
    u = rng.integers(0, 70_000, nnz).astype(np.int32)
    v = rng.integers(0, 5_000, nnz).astype(np.int32)
    r = rng.random(nnz).astype(np.float32)
    single = str(tmp_path / "single.bin")
    datagen.write_triplets(single, u, v, r)
    with open(single, "ab") as f:
        f.write(b"\x07" * 7)
    want_nodes, wm, wn = cs.read_problem(single)
    with cs.Context(0) as c:
        got_nnz, m, n = c.read_ratings(single)
        assert (got_nnz, m, n) == (nnz, wm, wn)
        gu, gv, gr = c.get_ratings()
        assert np.array_equal(gu, u) and np.array_equal(gv, v) and np.array_equal(gr.view(np.uint32), r.view(np.uint32))
        assert c.read_test(single) == nnz
    # shards of uneven sizes, one empty
    parts = []
    for i, k in enumerate((1000, 0, 4_200_000, 37)):
        pu = rng.integers(0, 900, k).astype(np.int32)
        pv = rng.integers(0, 300, k).astype(np.int32)
        pr = rng.random(k).astype(np.float32)
        datagen.write_triplets(str(tmp_path / f"sh{i}"), pu, pv, pr)
        parts.append((pu, pv, pr))
    want_nodes, wm, wn = cs.read_problem(str(tmp_path / "sh"))
    with cs.Context(0) as c:
        got_nnz, m, n = c.read_ratings(str(tmp_path / "sh"))
        assert (got_nnz, m, n) == (len(want_nodes), wm, wn)
        gu, gv, gr = c.get_ratings()
        assert np.array_equal(gu, want_nodes["u"]) and np.array_equal(gv, want_nodes["v"])
        assert np.array_equal(gr, want_nodes["rate"])
        # negative ids
        bad = str(tmp_path / "bad.bin")
        datagen.write_triplets(bad, np.array([0, -3], np.int32), np.array([1, 2], np.int32), np.ones(2, np.float32))
        with pytest.raises(cs.CumfError) as e:
            c.read_ratings(bad)
        assert e.value.code == 1 and "negative" in str(e.value)
        with pytest.raises(cs.CumfError) as e:
            c.read_ratings(str(tmp_path / "nope.bin"))
        assert e.value.code == 5


def test_streaming_reader_throughput(gpu, tmp_path):
    """The 99M-rating workload's training file (1.19 GB) from the page cache
    to the device: streaming reader vs read_problem + upload; both give the
    same arrays.  Rates are printed (this measures the host side)."""
    import time

    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    R = datagen.medium(seed=2, device="cpu")
    u, v, r = R.u.numpy(), R.v.numpy(), R.r.numpy()
    path = str(tmp_path / "medium.bin")
    datagen.write_triplets(path, u, v, r)
    del R
    with cs.Context(0) as c:
        c.read_ratings(path)          # warm the page cache and the runtime
        t0 = time.perf_counter()
        nnz, m, n = c.read_ratings(path)
        t_stream = time.perf_counter() - t0
        gu, gv, gr = c.get_ratings()
        assert np.array_equal(gu, u) and np.array_equal(gv, v) and np.array_equal(gr, r)
        t0 = time.perf_counter()
        nodes, m2, n2 = cs.read_problem(path)
        c.set_ratings_nodes(nodes, m2, n2)
        t_host = time.perf_counter() - t0
    gb = 12 * nnz / 1e9
    print(f"streaming reader {gb / t_stream:.2f} GB/s ({t_stream:.2f} s), read_problem + upload "
          f"{gb / t_host:.2f} GB/s ({t_host:.2f} s)")
    assert (m, n) == (m2, n2)


def test_cli_file_pipeline_equals_host_pipeline_in_serial_mode(gpu, tmp_path):
    """The CLI's 1x1 path (files streamed into HBM, device statistics) and
    cumf_sgd_train (host read_problem, host collect_data) give the same model
    file bit for bit in serial mode: integer ratings make the FP64 sums
    order-independent, so every later step sees the same numbers."""
    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    R = _data(nnz=40_000, n_test=4_000)
    u, v, r, tu, tv, tr = R.numpy()
    tr_path, te_path = str(tmp_path / "train.bin"), str(tmp_path / "test.bin")
    datagen.write_triplets(tr_path, u, v, r)
    datagen.write_triplets(te_path, tu, tv, tr)
    model = str(tmp_path / "model")
    out = subprocess.run([cs.CLI_PATH, "-k", "64", "-t", "3", "-l", "0.05", "-a", "0.08", "-b", "0.3", "--serial",
                          "--seed", "6", "-p", te_path, tr_path, model], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rm = [l.split()[-1] for l in out.stdout.splitlines() if "test RMSE" in l]
    nodes, m, n = cs.read_problem(tr_path)
    tnodes, _, _ = cs.read_problem(te_path)
    p = cs.make_param(k=64, num_iters=3, lambda_p=0.05, lambda_q=0.05, alpha=0.08, beta=0.3)
    P, Q, b, rmse = cs.sgd_train(p, nodes, m, n, precision=cs.FP16, seed=6, test=tnodes,
                                 opts=cs.make_opts(k=64, schedule=cs.SERIAL))
    host_model = str(tmp_path / "host_model")
    cs.save_model(host_model, P, Q, b)
    assert open(model, "rb").read() == open(host_model, "rb").read()
    assert rm == [f"{x:.4f}" for x in rmse]


def test_ranking_criteria_match_oracle_and_mf_predict(gpu, tmp_path):
    """Row/column MPR and AUC (mf-predict -e 10..13) on the GPU: equal to the
    oracle's restatement of calc_mpr_auc up to the FP64 summation order (the
    scores are the same sequential FP32 dots, so every comparison agrees), in
    every storage format's widened values, and printed by cumf_predict as the
    reference's compiled mf-predict prints them; duplicate positives are
    rejected."""
    import cumf_sgd_amd as cs
    import oracle
    from cumf_sgd_amd import datagen

    rng = np.random.default_rng(13)
    m, n, k = 3000, 900, 64
    P = rng.normal(0, 0.3, (m, k)).astype(np.float32)
    Q = rng.normal(0, 0.3, (n, k)).astype(np.float32)
    flat = rng.choice((m + 5) * (n + 4), 60_000, replace=False)
    tu = (flat // (n + 4)).astype(np.int32)
    tv = (flat % (n + 4)).astype(np.int32)
    tr = np.where(rng.random(len(tu)) < 0.5, 1.0, -1.0).astype(np.float32)
    b = np.float32(0.1)
    with cs.Context(0) as c:
        c.set_model(P, Q)
        c.set_test(tu, tv, tr)
        for crit in (10, 11, 12, 13):
            got = c.test_eval(crit, 1.0, b)
            want = oracle.evaluate(crit, tu, tv, tr, P, Q, b)
            assert abs(got - want) <= 1e-12 * abs(want), (crit, got, want)
        dup_u, dup_v = np.concatenate([tu, tu[:1]]), np.concatenate([tv, tv[:1]])
        c.set_test(dup_u, dup_v, np.concatenate([tr, np.ones(1, np.float32)]))
        if tr[0] > 0:
            with pytest.raises(cs.CumfError):
                c.test_eval(10, 1.0, b)
    # a trained FP16 model through the export scale (mult)
    u, v, r = tu[tr > 0] % m, tv[tr > 0] % n, np.ones(int((tr > 0).sum()), np.float32)
    with cs.Context(0) as c:
        c.set_ratings(u, v, r, m, n)
        c.init_factors(64, cs.FP16, seed=5)
        c.train(cs.make_opts(k=64, schedule=cs.HOGWILD, seed=5), 0, 2)
        c.set_test(tu, tv, tr)
        raw_p, _ = c.get_factors(0)
        raw_q, _ = c.get_factors(1)
        mult = np.float32(1.3)
        Pw = (oracle.widen(raw_p, None, 1) * mult).astype(np.float32)
        Qw = (oracle.widen(raw_q, None, 1) * mult).astype(np.float32)
        for crit in (10, 13):
            got = c.test_eval(crit, mult, b)
            want = oracle.evaluate(crit, tu, tv, tr, Pw, Qw, b)
            assert abs(got - want) <= 1e-12 * abs(want), (crit, got, want)
    model, test = str(tmp_path / "model"), str(tmp_path / "test.bin")
    cs.save_model(model, P, Q, b)
    datagen.write_triplets(test, tu, tv, tr)
    label = {10: "Row-wise MPR", 11: "Column-wise MPR", 12: "Row-wise AUC", 13: "Colmn-wise AUC"}
    for crit in (10, 11, 12, 13):
        out = subprocess.run([cs.PREDICT_PATH, "-e", str(crit), test, model], capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        line = out.stdout.strip().splitlines()[-1]
        assert line.startswith(label[crit] + " = "), line
        if oracle.have_mf_predict():
            assert line.split("=", 1)[1].strip() == oracle.mf_predict_eval(test, model, crit)


def test_c1_config_cli_beside_reference_cpu_trainer(gpu, tmp_path):
    """The small configuration as specified: 10,000 x 2,000, 900,000 / 100,000
    ratings, k=128, FP32 factors, lambda 0.05, alpha 0.08, beta 0.3, 20
    iterations, serialized, through the cumf_sgd command line
    (singleGPU/main.cpp's flags).  The model file scored by the reference's
    mf-predict gives the RMSE the command printed, within 0.005 of the
    serialized FP32 oracle (the command permutes the ids first, so the two
    runs visit the same ratings with other initial rows).  The reference's
    CPU trainer (test/mf-train: LIBMF's FPSGD, -s 1) runs beside it on the
    same files; its test RMSE and rate are printed (another optimizer, no
    bar)."""
    import time

    import cumf_sgd_amd as cs
    import oracle
    from cumf_sgd_amd import datagen

    R = datagen.small(seed=1)
    u, v, r, tu, tv, tr = R.numpy()
    tr_path, te_path = str(tmp_path / "train.bin"), str(tmp_path / "test.bin")
    datagen.write_triplets(tr_path, u, v, r)
    datagen.write_triplets(te_path, tu, tv, tr)
    model = str(tmp_path / "model")
    t0 = time.perf_counter()
    out = subprocess.run([cs.CLI_PATH, "-g", "0", "-l", "0.05", "-k", "128", "-t", "20", "-a", "0.08", "-b", "0.3",
                          "-s", "1", "--serial", "--precision", "fp32", "--seed", "1", "-p", te_path, tr_path, model],
                         capture_output=True, text=True, timeout=300)
    cli_s = time.perf_counter() - t0
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [l for l in out.stdout.splitlines() if "test RMSE" in l]
    assert len(lines) == 20
    cli_rmse = lines[-1].split()[-1]
    if oracle.have_mf_predict():
        assert oracle.mf_predict_rmse(te_path, model) == cli_rmse

    # the serialized FP32 oracle on the unpermuted ids
    std = np.float32(np.sqrt((r.astype(np.float64) ** 2).mean() - np.float64(r.astype(np.float64).mean()) ** 2))
    ave = np.float32(r.astype(np.float64).mean())
    scale = max(np.float32(1e-4), std)
    rs = (r * np.float32(1.0 / scale)).astype(np.float32)
    P, _ = oracle.init_factors(R.m, 128, 0, 1, 0)
    Q, _ = oracle.init_factors(R.n, 128, 0, 1, 1)
    lr = oracle.lr_table(0.08, 0.3, 0.0, 24)
    lam = np.float32(0.05) / scale
    P, Q, _, _ = oracle.sgd_serial(u, v, rs, P, Q, lr, 0, 20, k=128, G=64, prec=0, lambda_p=lam, lambda_q=lam,
                                   seed=1)
    fs = np.float32(np.sqrt(scale))
    want = oracle.rmse(tu, tv, tr, (P * fs).astype(np.float32), (Q * fs).astype(np.float32), ave / std)
    print(f"C1 cumf_sgd --serial fp32: test RMSE {cli_rmse} ({cli_s:.1f} s with file I/O); serialized oracle "
          f"{want:.4f}")
    assert abs(float(cli_rmse) - want) <= 0.005

    if os.access(oracle.MF_TRAIN, os.X_OK):
        ref = subprocess.run([oracle.MF_TRAIN, "-l2", "0.05", "-k", "128", "-t", "20", "-s", "1", "-p", te_path,
                              tr_path, str(tmp_path / "ref_model")], capture_output=True, text=True, timeout=300)
        assert ref.returncode == 0, ref.stdout + ref.stderr
        rows = [l.split() for l in ref.stdout.splitlines() if l.split() and l.split()[0].isdigit()]
        perf = [l for l in ref.stdout.splitlines() if l.startswith("perf:")]
        va = float(rows[-1][2])
        print(f"C1 test/mf-train -s 1: test RMSE {va:.4f} after {len(rows)} iterations, "
              f"{perf[0] if perf else 'no perf line'}")
        assert len(rows) == 20 and np.isfinite(va)


def test_cli_default_schedule_on_tiny_and_small_inputs(gpu, tmp_path):
    """The command line's 1x1 default on a 7-rating file with fewer ids than
    strata and on a small problem whose rows fit in the L2s (the default is
    then the strata with four passes over the diagonals per epoch): both
    train and write the model file; on the small problem the default ends
    no worse than the stratum schedule with one pass (--schedule stratum) or
    two passes (--rounds 2, which implies the strata)."""
    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    tiny = str(tmp_path / "tiny.bin")
    datagen.write_triplets(tiny, np.array([0, 1, 2, 0, 1, 2, 1], np.int32), np.array([0, 0, 1, 1, 2, 2, 0], np.int32),
                           np.array([1, 2, 3, 4, 5, 4, 3], np.float32))
    out = subprocess.run([cs.CLI_PATH, "-k", "32", "-t", "3", "-s", "8", "--seed", "2", tiny, str(tmp_path / "tm")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert os.path.getsize(str(tmp_path / "tm")) == 20 + 4 * (3 + 3) * 32

    R = _data(nnz=150_000, n_test=15_000)
    u, v, r, tu, tv, tr = R.numpy()
    tr_path, te_path = str(tmp_path / "train.bin"), str(tmp_path / "test.bin")
    datagen.write_triplets(tr_path, u, v, r)
    datagen.write_triplets(te_path, tu, tv, tr)
    last = {}
    for sched in ("default", "stratum", "rounds"):
        extra = {"default": [], "stratum": ["--schedule", "stratum"], "rounds": ["--rounds", "2"]}[sched]
        out = subprocess.run([cs.CLI_PATH, "-l", "0.05", "-a", "0.08", "-b", "0.3", "-s", "256", "-k", "64", "-t", "8",
                              "--seed", "4", *extra, "-p", te_path, tr_path, str(tmp_path / f"m_{sched}")],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        last[sched] = float([l for l in out.stdout.splitlines() if "test RMSE" in l][-1].split()[-1])
    print(f"default (strata, 4 passes) {last['default']:.4f}, stratum {last['stratum']:.4f}, "
          f"two passes {last['rounds']:.4f}")
    assert last["default"] <= last["stratum"] + 0.005 and last["default"] <= last["rounds"] + 0.005
