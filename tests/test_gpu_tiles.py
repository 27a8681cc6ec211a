"""The tile schedule (CUMF_SCHED_TILE) through the C ABI: the tile partition
against a numpy restatement, tile epochs bit for bit against the CPU oracle
wherever they are deterministic, graph replay, refusals and convergence.

A tile is a range of consecutive rows of one row band of the s x s strata;
its records (within one block) are trained by one workgroup that holds the
tile's P rows in LDS for the launch.  With one worker per tile and one tile
per XCD in flight (tiles_in_flight = 1) no two concurrent workers share a
row, so an epoch is deterministic: the diagonals in order, and within a
block its tiles largest first, each tile's windows drawn with the tile's
index in the tile list as tag.
"""
import numpy as np
import pytest
from ranges_rule import capped_nearest_ranges, first_crossing_ranges, least_largest_range

pytestmark = pytest.mark.gpu


def _problem(seed, m, n, nnz, n_test=0):
    from cumf_sgd_amd import datagen

    R = datagen.uniform(seed, m, n, nnz, n_test)
    return R.numpy() + (m, n)


def _ctx(u, v, r, m, n):
    import cumf_sgd_amd as cs

    c = cs.Context(0)
    c.set_ratings(u, v, r, m, n)
    return c


def _row_key(row0, rows, m):
    """Tile of every row: the first row of the tile interval holding it."""
    key = np.full(m, -1, np.int64)
    for a, b in zip(row0, rows):
        key[a:a + b] = a
    return key


@pytest.mark.parametrize("s,tile_rows", [(8, 16), (8, 0), (4, 7), (16, 64)])
def test_tile_partition_is_strata_then_stable_tile_column_sort(gpu, s, tile_rows):
    """The tile partition has the blocks of cumf_partition_strata; inside a
    block the records are the stratum block's, stably sorted by (tile of
    the row, column).  Tiles cover each block exactly, hold at most
    tile_rows rows of one row band, and every record's row lies in its
    tile."""
    u, v, r, _, _, _, m, n = _problem(seed=21, m=900, n=300, nnz=40_000)
    with _ctx(u, v, r, m, n) as c:
        off_s = c.partition_strata(s)
        su, sv, sr = c.get_ratings()
    with _ctx(u, v, r, m, n) as c:
        off = c.partition_tiles(s, tile_rows)
        gu, gv, gr = c.get_ratings()
        toff, row0, rows = c.get_tiles()
    assert np.array_equal(off, off_s)
    assert toff[0] == 0 and toff[-1] == len(u) and np.all(np.diff(toff) > 0)
    cap = tile_rows or 96   # cumf_tile_rows(128, FP16) before factors exist
    assert rows.max() <= cap and rows.min() >= 1
    # every block is a whole number of tiles
    assert set(off.tolist()) <= set(toff.tolist())
    key = _row_key(row0, rows, m)
    tile_of_rec = np.repeat(np.arange(len(row0)), np.diff(toff))
    assert np.all(gu >= row0[tile_of_rec]) and np.all(gu < row0[tile_of_rec] + rows[tile_of_rec])
    # row ranges of distinct tiles never overlap
    iv = sorted(set(zip(row0.tolist(), rows.tolist())))
    assert all(a0 + a1 <= b0 for (a0, a1), (b0, _) in zip(iv, iv[1:]))
    # expected order: each stratum block stably sorted by (tile, column)
    blk = np.repeat(np.arange(len(off) - 1), np.diff(off))
    order = np.lexsort((sv, key[su], blk))
    assert np.array_equal(gu, su[order]) and np.array_equal(gv, sv[order]) and np.array_equal(gr, sr[order])


def _tile_windows(n, W, S, Lmax, D, R):
    """A tile's window length and window count, restated from the planner:
    N = ceil(n / L) windows, pass r running windows [N r / R, N (r+1) / R)
    on w workers of u windows (u = ceil(n_r / W), w = ceil(n_r / u)); L is
    the multiple of D up to Lmax whose passes issue the fewest slot-steps
    (whole waves of S slots, a window 16 steps longer than its records),
    ties to fewer samples, then to the longer window."""
    best = None
    for L in range(Lmax, D - 1, -D):
        N = -(-n // L)
        cost = samples = 0
        for r in range(R):
            nr = N * (r + 1) // R - N * r // R
            if nr == 0:
                continue
            u = -(-nr // W)
            w = -(-nr // u)
            cost += -(-w // S) * S * u * (L + 16)
            samples += w * u * L
        if best is None or (cost, samples) < best[:2]:
            best = (cost, samples, L, N)
    return best[2], best[3]


def test_tile_windows_restatement_fills_waves():
    """The restated rule on the shapes it is for: a tile of 2388 records on
    32 workers of 8-slot waves takes 26-record windows (92 windows, 31
    workers of 3 windows, 32 slots issued) instead of 32-record ones (25
    workers of 3 windows: 7 idle slots); one worker per tile takes the
    length with the least over-sampling per window step."""
    assert _tile_windows(2388, 32, 8, 32, 2, 1) == (26, 92)
    assert _tile_windows(20, 1, 8, 32, 2, 1) == (20, 1)
    assert _tile_windows(64, 1, 8, 32, 2, 1) == (32, 2)
    assert _tile_windows(0, 32, 8, 32, 2, 2) == (32, 0)


def _oracle_tiles(gu, gv, gr, off, toff, s, P, Q, ps, qs, epochs, rounds, prec, k, G, Lmax, lp, lq, seed):
    """The deterministic tile epoch restated: passes, diagonals, each block's
    tiles largest first (stable), a tile's windows of the pass in order, one
    worker per tile with the tile's own window length."""
    import oracle

    nt = len(toff) - 1
    lr = oracle.lr_table(0.08, 0.3, 0.0, epochs + 4)
    tile_blk = np.searchsorted(off, toff[:-1], side="right") - 1
    for ep in range(epochs):
        nxt = np.zeros(nt, np.int64)
        for rr in range(rounds):
            for d in range(s):
                for i in range(s):
                    b = i * s + (i + d) % s
                    ts = np.flatnonzero(tile_blk == b)
                    sizes = toff[ts + 1] - toff[ts]
                    for t in ts[np.argsort(-sizes, kind="stable")]:
                        lo, hi = int(toff[t]), int(toff[t + 1])
                        L, n_all = _tile_windows(hi - lo, 1, 64 // G, Lmax, 2, rounds)
                        n_pass = n_all * (rr + 1) // rounds - n_all * rr // rounds
                        if n_pass == 0:
                            continue
                        oracle.sgd_block_windows(gu[lo:hi], gv[lo:hi], gr[lo:hi], P, Q, lr, ep, int(t), int(nxt[t]),
                                                 int(n_pass), k=k, G=G, prec=prec, lambda_p=lp, lambda_q=lq,
                                                 seed=seed, batch_len=L, ps=ps, qs=qs)
                        nxt[t] += n_pass
        if prec == 3:
            P, ps = oracle.fp8_rescale(P, ps)
            Q, qs = oracle.fp8_rescale(Q, qs)
    return P, Q, ps, qs


@pytest.mark.parametrize("prec,tile_rows,rounds,k", [
    (0, 0, 1, 64), (1, 0, 1, 64), (1, 8, 1, 64), (1, 8, 2, 64), (0, 8, 1, 128), (3, 8, 1, 64), (4, 8, 2, 64),
    (2, 8, 1, 32), (1, 5, 1, 128)])
def test_tile_epochs_match_oracle_when_deterministic(gpu, prec, tile_rows, rounds, k):
    """One worker per tile and one tile per XCD in flight: GPU factors after
    three epochs equal the oracle's bit for bit (P rows trained in LDS, Q
    rows in memory), with one tile per block (tile_rows 0: the default,
    larger than a band) and with several tiles per block, one and two passes
    over the diagonals, every storage format."""
    import cumf_sgd_amd as cs

    u, v, r, _, _, _, m, n = _problem(seed=22, m=320, n=200, nnz=12_000)
    s, seed, lp, lq, L = 8, 5, 0.05, 0.04, 32
    with _ctx(u, v, r, m, n) as c:
        c.init_factors(k, prec, seed=seed)
        off = c.partition_tiles(s, tile_rows)
        gu, gv, gr = c.get_ratings()
        toff, row0, rows = c.get_tiles()
        if tile_rows == 0:
            assert len(toff) - 1 <= s * s   # one tile per non-empty block
        P, ps = c.get_factors(0)
        Q, qs = c.get_factors(1)
        c.train(cs.make_opts(k=k, schedule=cs.TILE, num_workers=1, tiles_in_flight=1, seed=seed, lambda_p=lp,
                             lambda_q=lq, stratum_rounds=rounds, batch_len=L), 0, 3)
        G = c.last_launch().lanes_per_sample
        P1, ps1 = c.get_factors(0)
        Q1, qs1 = c.get_factors(1)
    P, Q = np.ascontiguousarray(P).copy(), np.ascontiguousarray(Q).copy()
    P, Q, ps, qs = _oracle_tiles(gu, gv, gr, off, toff, s, P, Q, ps, qs, 3, rounds, prec, k, G, L, lp, lq, seed)
    assert np.array_equal(P1.view(np.uint8), P.view(np.uint8))
    assert np.array_equal(Q1.view(np.uint8), Q.view(np.uint8))
    if prec == 3:
        assert np.array_equal(ps1, ps) and np.array_equal(qs1, qs)


@pytest.mark.parametrize("prec", [1, 3])
def test_tile_graph_replay_equals_eager(gpu, prec):
    """Three tile epochs replayed from a captured graph equal the eager
    launches bit for bit (deterministic configuration, eager work between
    replays)."""
    import cumf_sgd_amd as cs

    u, v, r, _, _, _, m, n = _problem(seed=23, m=400, n=160, nnz=10_000)
    out = []
    for graph in (0, 1):
        with _ctx(u, v, r, m, n) as c:
            c.partition_tiles(8, 12)
            c.init_factors(64, prec, seed=2)
            opts = cs.make_opts(k=64, schedule=cs.TILE, num_workers=1, tiles_in_flight=1, seed=2, batch_len=32,
                                use_graph=graph)
            for ep in range(3):
                c.train(opts, ep, ep + 1)
                c.get_factors(0)   # eager work between replays
            out.append((c.get_factors(0), c.get_factors(1)))
    (pe, qe), (pg, qg) = out
    assert np.array_equal(pe[0].view(np.uint8), pg[0].view(np.uint8))
    assert np.array_equal(qe[0].view(np.uint8), qg[0].view(np.uint8))


def test_tile_full_workgroups_train_every_rating_window(gpu):
    """The default tile launch (a workgroup of workers per tile, all tiles of
    a diagonal at once) trains the planned updates, leaves the factors
    finite and lowers the training error of a small problem."""
    import cumf_sgd_amd as cs

    u, v, r, tu, tv, tr, m, n = _problem(seed=24, m=3000, n=800, nnz=200_000, n_test=20_000)
    with _ctx(u, v, r, m, n) as c:
        c.set_test(tu, tv, tr)
        c.partition_tiles(8)
        c.init_factors(128, cs.FP16, seed=3)
        r0 = c.test_rmse(1.0, 0.0)
        opts = cs.make_opts(k=128, schedule=cs.TILE, seed=3, batch_len=32)
        c.train(opts, 0, 5)
        info = c.last_launch()
        toff, _, _ = c.get_tiles()
        Lmax, W = info.batch_len, info.num_workers
        assert W == 32 and info.lanes_per_sample == 8   # 32 bytes of an FP16 row per lane
        planned = 0
        for size in np.diff(toff):
            L, N = _tile_windows(int(size), W, 8, Lmax, 2, 1)
            uc = -(-N // W)        # fewest windows per worker
            w = -(-N // uc)        # then fewest workers keeping them
            assert w * uc * L >= size
            planned += w * uc * L
        assert info.updates_per_epoch == planned
        r1 = c.test_rmse(1.0, 0.0)
        P, _ = c.get_factors(0)
        assert np.all(np.isfinite(P.view(np.float16).astype(np.float32)))
        assert r1 < r0


def test_tile_refusals(gpu):
    """The tile schedule needs a tile partition, the default lane group,
    depth 2, no coherent mode, at most 16 waves of workers per tile, windows
    of at most 32 records and tiles that fit in LDS; a later partition drops
    the tiles."""
    import cumf_sgd_amd as cs

    u, v, r, _, _, _, m, n = _problem(seed=25, m=500, n=200, nnz=5_000)
    with _ctx(u, v, r, m, n) as c:
        c.init_factors(128, cs.FP16, seed=1)
        base = dict(k=128, schedule=cs.TILE, seed=1, batch_len=32)
        with pytest.raises(cs.CumfError) as e:
            c.train(cs.make_opts(**base), 0, 1)
        assert e.value.code == 3
        c.partition_tiles(8)
        for bad in (dict(lanes_per_sample=64), dict(coherent=1), dict(pipeline_depth=4), dict(num_workers=1000),
                    dict(tiles_in_flight=-1), dict(batch_len=64)):
            with pytest.raises(cs.CumfError) as e:
                c.train(cs.make_opts(**{**base, **bad}), 0, 1)
            assert e.value.code == 1, bad
        c.train(cs.make_opts(**base), 0, 1)   # still usable after the refusals
        c.train(cs.make_opts(**{**base, "batch_len": 0}), 0, 1)
        assert c.last_launch().batch_len == 32   # the tile schedule's longest window by default
        c.partition_tiles(8, 700)   # 700 rows of 256 B exceed the LDS with the staged records
        with pytest.raises(cs.CumfError) as e:
            c.train(cs.make_opts(**base), 0, 1)
        assert e.value.code == 1
        c.partition_strata(8)
        with pytest.raises(cs.CumfError) as e:
            c.train(cs.make_opts(**base), 0, 1)
        assert e.value.code == 3


def test_epoch_per_call_reuses_the_plan_bit_exactly(gpu):
    """One epoch per cumf_train call reuses the task table planned by the
    call before (the plan key matches), also after a call that planned
    another schedule in between (a Hogwild epoch with a zero learning rate,
    which leaves the factors as they are): three such calls equal one call
    of three epochs bit for bit, for the tile and the stratum schedule."""
    import cumf_sgd_amd as cs

    u, v, r, _, _, _, m, n = _problem(seed=26, m=320, n=200, nnz=12_000)
    for sched in (cs.TILE, cs.STRATUM):
        out = []
        for per_call in (False, True):
            with _ctx(u, v, r, m, n) as c:
                c.init_factors(64, cs.FP16, seed=4)
                if sched == cs.TILE:
                    c.partition_tiles(8, 8)
                else:
                    c.partition_strata(8)
                opts = cs.make_opts(k=64, schedule=sched, num_workers=1, tiles_in_flight=1, seed=4, batch_len=32,
                                    use_graph=1)
                idle = cs.make_opts(k=64, schedule=cs.HOGWILD, seed=4, alpha=0.0, beta=0.0)
                if per_call:
                    for ep in range(3):
                        c.train(opts, ep, ep + 1)
                        if ep == 1:
                            c.train(idle, 0, 1)
                else:
                    c.train(opts, 0, 3)
                out.append((c.get_factors(0)[0], c.get_factors(1)[0]))
        (p3, q3), (p1, q1) = out
        assert np.array_equal(p3.view(np.uint8), p1.view(np.uint8)), sched
        assert np.array_equal(q3.view(np.uint8), q1.view(np.uint8)), sched


def test_strata_ranges_balance_the_largest_when_each_xcd_trains_one_block(gpu):
    """cumf_partition_strata's row and column ranges, restated from the
    degree histograms of the device's ids: for s <= 8 (one block per XCD and
    diagonal) the largest range holds exactly the least count any contiguous
    cut allows, each range closed nearest the remaining even share; for
    s > 8 a range ends at the first row reaching the next 1/s.  The columns
    are Zipf-skewed, so one column holds more than 1/16 of the ratings."""
    rng = np.random.default_rng(31)
    m, n, nnz = 3000, 400, 80_000
    p = 1.0 / np.arange(1, n + 1) ** 1.1
    v = rng.choice(n, nnz, p=p / p.sum()).astype(np.int32)
    u = rng.integers(0, m, nnz).astype(np.int32)
    r = rng.uniform(1, 5, nnz).astype(np.float32)
    with _ctx(u, v, r, m, n) as c:
        for s in (2, 4, 8, 16, 24):
            off = np.asarray(c.partition_strata(s))
            gu, gv, _ = c.get_ratings()
            blocks = np.diff(off).reshape(s, s)
            for ids, size, bands in ((gu, m, blocks.sum(1)), (gv, n, blocks.sum(0))):
                deg = np.bincount(np.asarray(ids), minlength=size).astype(np.int64)
                want = capped_nearest_ranges(deg, s) if s <= 8 else first_crossing_ranges(deg, s)
                assert np.array_equal(bands, want), (s, bands, want)
                if s <= 8:
                    assert bands.max() == least_largest_range(deg, s)
        hot = np.bincount(np.asarray(gv), minlength=n).max()
        assert hot * 16 > nnz


@pytest.mark.parametrize("deg,s", [([1, 3, 1, 0, 5, 1], 3), ([0, 0, 5, 1, 1, 1, 3, 3, 8, 8, 1], 5),
                                   ([30, 5, 0, 1, 13, 13, 5, 3, 30, 8, 5], 4), ([3, 13, 3, 1, 30, 13, 1, 30, 1], 5)])
def test_strata_ranges_stay_within_the_least_largest_range(gpu, deg, s):
    """Degree histograms where closing every range at the boundary nearer
    the even share would leave a later range above the least largest count:
    the library's row and column ranges (both sides given the histogram)
    equal the restated cut and their largest is that least count."""
    deg = np.array(deg, dtype=np.int64)
    m = len(deg)
    rng = np.random.default_rng(5)
    u = np.repeat(np.arange(m, dtype=np.int32), deg)
    v = rng.permutation(u).astype(np.int32)
    r = rng.uniform(1, 5, len(u)).astype(np.float32)
    with _ctx(u, v, r, m, m) as c:
        off = np.asarray(c.partition_strata(s))
        blocks = np.diff(off).reshape(s, s)
    want = capped_nearest_ranges(deg, s)
    assert np.array_equal(blocks.sum(1), want) and np.array_equal(blocks.sum(0), want)
    assert want.max() == least_largest_range(deg, s)
