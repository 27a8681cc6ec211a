"""Pins of the CPU oracle against facts fixed outside it: published known-answer
vectors, independent library implementations of the number formats, the
reference's own compiled evaluator, closed forms and brute force."""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# ---------------------------------------------------------------- Philox
# Known-answer vectors of Philox4x32-10 published with Random123 (kat_vectors).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert list(oracle.philox(ctr, key)) == want


def test_window_starts_uniform_and_in_range():
    nnz = 1_000_003
    s = np.array([oracle.window_start(9, 0, 0, w, 0, b, nnz) for w in range(20) for b in range(500)])
    assert s.min() >= 0 and s.max() < nnz
    # uniform over [0, nnz): mean and quartiles within a few standard errors
    assert abs(s.mean() / nnz - 0.5) < 0.01
    assert abs(np.mean(s < nnz // 4) - 0.25) < 0.015
    # every record reachable: starts are not on a coarse lattice (FP32 product in the reference)
    big = 3_101_144_313
    t = np.array([oracle.window_start(1, 0, 0, 0, 0, b, big) for b in range(2000)])
    assert np.any(t % 256 != 0) and t.max() < big


# ---------------------------------------------------------------- formats
def _probe_floats(n=200_000, seed=0):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    return f[np.isfinite(f)]


def test_f16_matches_numpy_rne():
    f = _probe_floats()
    f = np.concatenate([f, np.float32([65504, 65519.99, 65520, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, -0.0, 0.0])])
    with np.errstate(over="ignore"):   # values past FP16's range round to infinity, on purpose
        want = f.astype(np.float16).view(np.uint16)
    got = np.array([oracle.f32_to_f16(float(x)) for x in f], np.uint16)
    assert np.array_equal(got, want)
    h = np.arange(0, 65536, dtype=np.uint32).astype(np.uint16)
    back = np.array([oracle.f16_to_f32(int(x)) for x in h], np.float32)
    ref = h.view(np.float16).astype(np.float32)
    same = (back.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(back) & np.isnan(ref))
    assert same.all()


def test_bf16_matches_torch_rne():
    import torch

    f = _probe_floats(100_000, 1)
    want = torch.from_numpy(f).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([oracle.f32_to_bf16(float(x)) for x in f], np.uint16)
    assert np.array_equal(got, want)


def test_e4m3_matches_torch_in_range():
    import torch

    rng = np.random.default_rng(2)
    f = np.concatenate([rng.uniform(-447.9, 447.9, 50_000), rng.normal(0, 0.01, 50_000),
                        rng.normal(0, 1e-3, 20_000)]).astype(np.float32)
    want = torch.from_numpy(f).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    got = np.array([oracle.f32_to_e4m3_rne(float(x)) for x in f], np.uint8)
    assert np.array_equal(got, want)
    allb = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    back = np.array([oracle.e4m3_to_f32(i) for i in range(256)], np.float32)
    assert np.array_equal(np.isnan(back), np.isnan(allb))
    ok = ~np.isnan(allb)
    assert np.array_equal(back[ok], allb[ok])
    # saturation instead of nan above the largest finite value
    assert oracle.f32_to_e4m3_rne(1e6) == 0x7E and oracle.f32_to_e4m3_rne(-500.0) == 0xFE


def test_e4m3_stochastic_rounding_brackets_and_is_unbiased():
    grid = np.array([oracle.e4m3_to_f32(i) for i in range(0x7F)], np.float32)
    rng = np.random.default_rng(3)
    for x in [0.3, 1.1, 3.3e-3, 1.7e-3, 250.0, 0.01171875]:
        lo = grid[grid <= x].max()
        hi = grid[grid >= x].min()
        vals = np.array([oracle.e4m3_to_f32(oracle.f32_to_e4m3_sr(x, int(r)))
                         for r in rng.integers(0, 65536, 40_000)])
        assert set(np.unique(vals)) <= {lo, hi}
        se = (hi - lo) / 2 / np.sqrt(len(vals))
        assert abs(vals.mean() - np.float32(x)) < 5 * se + 1e-12
    # representable values are fixed points for every random draw
    for b in [0x01, 0x08, 0x3C, 0x7E]:
        x = oracle.e4m3_to_f32(b)
        assert all(oracle.f32_to_e4m3_sr(x, r) == b for r in (0, 1, 0x7FFF, 0xFFFF))


def test_bf16_stochastic_rounding_counts_exactly():
    """BF16-SR over all 65536 random words: the result is the truncation or
    the next BF16 away from zero, and it rounds up for exactly (low 16 bits
    of x) words, i.e. with probability equal to the dropped fraction.
    Truncation itself is torch's bfloat16 of the value with its low bits
    cleared; representable values are fixed points; NaN stays NaN."""
    import torch

    rng = np.random.default_rng(9)
    xs = np.concatenate([rng.normal(0, 3, 6).astype(np.float32), np.float32([1e-30, -2.5e-3, 3.0e38])])
    for x in xs:
        bits = int(np.float32(x).view(np.uint32))
        trunc = bits >> 16
        want_trunc = int(torch.tensor(np.uint32(bits & 0xFFFF0000).view(np.float32)).to(torch.bfloat16)
                         .view(torch.int16).item()) & 0xFFFF
        assert want_trunc == trunc
        outs = np.array([oracle.f32_to_bf16_sr(float(x), r) for r in range(65536)])
        assert set(np.unique(outs)) <= {trunc, trunc + 1}
        assert int((outs == trunc + 1).sum()) == bits & 0xFFFF
    for b in (0x3F80, 0x0001, 0xC2F7, 0x7F7F):
        x = oracle.bf16_to_f32(b)
        assert all(oracle.f32_to_bf16_sr(x, r) == b for r in (0, 1, 0x8000, 0xFFFF))
    assert oracle.f32_to_bf16_sr(float("nan"), 0xFFFF) & 0x7FC0 == 0x7FC0
    assert oracle.f32_to_bf16_sr(float("inf"), 0xFFFF) == 0x7F80


def test_e4m3_stochastic_rounding_matches_gfx950_instruction_vectors():
    """The oracle's SR rule against 21,302 outputs of v_cvt_sr_fp8_f32 (the
    conversion the update kernel stores FP8 with), written on a gfx950 GPU by
    tests/golden/make_e4m3_sr_vectors.py: subnormal-range values at and around
    the round-up threshold of their 20-bit fraction, normal values at the
    threshold, random values of both signs, saturation, zeros and NaNs."""
    d = np.load(os.path.join(GOLDEN, "e4m3_sr_hw.npz"))
    f, r16, want = d["f"], d["r16"], d["byte"]
    got = np.array([oracle.f32_to_e4m3_sr(float(x), int(r)) for x, r in zip(f, r16)], np.uint8)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(float(f[i]), int(r16[i]), int(want[i]), int(got[i])) for i in bad[:5]]
    # the vectors reach every regime the rule distinguishes
    ax = np.abs(f)
    assert (ax < 2.0 ** -6).sum() > 1000 and ((ax >= 2.0 ** -6) & (ax < 448)).sum() > 1000
    assert (want == 0xFF).any() and (want == 0x7E).any() and (want == 0xFE).any()


@pytest.mark.parametrize("amax,target", [(0.0884, 14.0), (3.0, 56.0), (56.0, 56.0), (56.01, 56.0), (1e-7, 14.0)])
def test_pow2_scale_is_smallest_power_of_two(amax, target):
    s = oracle.pow2_scale(amax, target)
    e = math.log2(s)
    assert e == int(e)
    assert np.float32(amax) <= np.float32(target) * s
    assert np.float32(amax) > np.float32(target) * (s / 2)


def test_pow2_scale_stays_normal():
    """Degenerate rows (|row| near 0) get 2^-126, never a subnormal or zero
    scale; the largest finite amax stays below the 2^126 cap."""
    assert oracle.pow2_scale(1e-38, 56.0) == 2.0 ** -126
    assert oracle.pow2_scale(float(np.finfo(np.float32).tiny) / 8, 14.0) == 2.0 ** -126
    s = oracle.pow2_scale(float(np.finfo(np.float32).max), 14.0)
    assert s == 2.0 ** 125 and np.isfinite(np.float32(1.0) / np.float32(s))


# ---------------------------------------------------------------- schedule
def test_lr_table_known_values():
    # alpha=0.08, beta=0.3: 0.08, 0.0615385, 0.0432777, 0.0312641, 0.0235294
    lr = oracle.lr_table(0.08, 0.3, 0.0, 5)
    assert np.allclose(lr, [0.08, 0.0615385, 0.0432777, 0.0312641, 0.0235294], rtol=0, atol=5e-8)
    assert lr[0] == np.float32(np.float32(0.08))
    lr2 = oracle.lr_table(0.08, 0.3, 0.01, 3)
    assert np.allclose(lr2 - lr[:3], 0.01, atol=1e-7)


# ---------------------------------------------------------------- host preprocessing
def test_random_map_matches_libstdcxx_random_shuffle():
    """The restatement (rand() % (i+1) swaps) equals std::random_shuffle as
    compiled into the product library."""
    import cumf_sgd_amd as cs

    for seed, size in [(0, 1), (1, 2), (7, 1000), (12345, 17770), (2, 480_189)]:
        a = oracle.gen_random_map(seed, size)
        assert np.array_equal(a, cs.gen_random_map(seed, size))
        assert np.array_equal(np.sort(a), np.arange(size))
        inv = oracle.gen_inv_map(a)
        assert np.array_equal(inv[a], np.arange(size))


def _nodes(u, v, r):
    x = np.zeros(len(u), oracle.NODE_DTYPE)
    x["u"], x["v"], x["rate"] = u, v, r
    return x


def test_grid_problem_hand_example():
    # 2x2 grid over m=4, n=4 (u_seg = v_seg = 2): blocks 0:(u<2,v<2) 1:(u<2,v>=2) 2 3
    R = _nodes([3, 0, 1, 2, 0, 3], [3, 0, 2, 1, 1, 0], [1, 2, 3, 4, 5, 6])
    out, gs, off = oracle.grid_problem(R, 4, 4, 2, 2)
    assert list(gs) == [2, 1, 2, 1] and list(off) == [0, 2, 3, 5, 6]
    # cycle-swap trace by hand:
    # start [(3,3),(0,0),(1,2),(2,1),(0,1),(3,0)], pivots [0,2,3,5]
    # g0 p0:(3,3)->b3 swap with 5 -> [(3,0),...,(3,3)] piv3=6; p0:(3,0)->b2 swap with 3:
    #   [(2,1),(0,0),(1,2),(3,0),(0,1),(3,3)] piv2=4; p0:(2,1)->b2 swap with 4:
    #   [(0,1),(0,0),(1,2),(3,0),(2,1),(3,3)] piv2=5; p0 ok, p1 ok
    assert [(int(a), int(b)) for a, b in zip(out["u"], out["v"])] == [(0, 1), (0, 0), (1, 2), (3, 0), (2, 1), (3, 3)]


def test_grid_problem_invariants_and_product_equality():
    import cumf_sgd_amd as cs

    rng = np.random.default_rng(4)
    m, n = 1000, 300
    R = _nodes(rng.integers(0, m, 20000), rng.integers(0, n, 20000), rng.random(20000).astype(np.float32))
    for ux, vy in [(1, 1), (3, 2), (7, 5), (16, 16)]:
        out, gs, off = oracle.grid_problem(R, m, n, ux, vy)
        us, vs = oracle.grid_segments(m, n, ux, vy)
        bid = (out["u"] // us) * vy + out["v"] // vs
        assert np.all(np.diff(bid) >= 0)
        assert np.array_equal(np.bincount(bid, minlength=ux * vy), gs)
        assert np.array_equal(np.sort(out, order=["u", "v", "rate"]), np.sort(R, order=["u", "v", "rate"]))
        mine = R.copy()
        gs2, off2, us2, vs2 = cs.grid_problem(mine, m, n, ux, vy)
        assert np.array_equal(mine, out) and np.array_equal(off2, off) and (us2, vs2) == (us, vs)


def test_collect_data_matches_exact_sums():
    rng = np.random.default_rng(5)
    r = rng.integers(1, 6, 100_000).astype(np.float32)
    R = _nodes(np.zeros(len(r)), np.zeros(len(r)), r)
    ave, std = oracle.collect_data(R)
    ex = math.fsum(r.astype(np.float64)) / len(r)
    ex2 = math.fsum(r.astype(np.float64) ** 2) / len(r)
    assert ave == np.float32(ex) and std == np.float32(math.sqrt(ex2 - ex * ex))
    import cumf_sgd_amd as cs

    assert cs.collect_data(R) == (ave, std)


def test_scale_problem_localises_ids():
    R = _nodes([5, 9, 2], [1, 7, 3], [2.0, 4.0, 5.0])
    out = oracle.scale_problem(R, np.float32(0.5), 2, 2, 5, 4)
    assert list(out["u"]) == [0, 4, 2] and list(out["v"]) == [1, 3, 3]
    assert list(out["rate"]) == [1.0, 2.0, 2.5]


def test_inverse_shuffle_is_the_gather_by_forward_map():
    rng = np.random.default_rng(6)
    for size in [1, 2, 17, 1000]:
        pm = rng.permutation(size).astype(np.int32)
        inv = oracle.gen_inv_map(pm)
        vec = rng.random((size, 3)).astype(np.float32)
        assert np.array_equal(oracle.inv_shuffle(vec, inv), vec[pm])


# ---------------------------------------------------------------- SGD arithmetic
def test_tree_dot_is_exact_on_integers_and_order_sensitive_otherwise():
    rng = np.random.default_rng(7)
    p = rng.integers(-8, 8, 128).astype(np.float32)
    q = rng.integers(-8, 8, 128).astype(np.float32)
    for G in (64, 32, 16, 8, 4):
        assert oracle.tree_dot(p, q, G) == float(np.dot(p.astype(np.float64), q.astype(np.float64)))
    p = rng.normal(size=128).astype(np.float32)
    q = rng.normal(size=128).astype(np.float32)
    exact = math.fsum(float(a) * float(b) for a, b in zip(p, q))
    for G in (64, 16):
        assert abs(oracle.tree_dot(p, q, G) - exact) < 1e-5


@pytest.mark.parametrize("k,G", [(128, 8), (128, 16), (64, 8), (32, 4)])
def test_split_lane_layout_covers_half_rows_and_reorders_the_dot(k, G):
    """32-byte lanes: lane l holds elements [l h, (l+1) h) of each half of the
    row (h = k / G / 2).  The layout is a bijection, each piece index covers
    one half row contiguously across the lanes, and the dot product in it
    equals the contiguous-layout dot product of the rows permuted into that
    order (the permutation written out here from the definition)."""
    E = k // G
    h = E // 2
    got = np.array([[oracle.lane_elem(k, G, h, l, i) for i in range(E)] for l in range(G)])
    want = np.array([[(i // h) * (k // 2) + l * h + i % h for i in range(E)] for l in range(G)])
    assert np.array_equal(got, want)
    assert sorted(got.ravel().tolist()) == list(range(k))
    for c in range(2):
        piece = got[:, c * h:(c + 1) * h]
        assert np.array_equal(np.sort(piece.ravel()), np.arange(c * k // 2, (c + 1) * k // 2))
        assert np.all(np.diff(piece, axis=1) == 1)
    rng = np.random.default_rng(k + G)
    p = rng.normal(size=k).astype(np.float32)
    q = rng.normal(size=k).astype(np.float32)
    perm = want.ravel()
    assert oracle.tree_dot_split(p, q, G, h) == oracle.tree_dot(p[perm], q[perm], G)
    pi = rng.integers(-8, 8, k).astype(np.float32)
    qi = rng.integers(-8, 8, k).astype(np.float32)
    assert oracle.tree_dot_split(pi, qi, G, h) == float(np.dot(pi.astype(np.float64), qi.astype(np.float64)))


def test_single_update_known_answer_and_gradient():
    """k=2 hand-computed update, then a finite-difference gradient check of the
    loss 1/2 e^2 + 1/2 lp|p|^2 + 1/2 lq|q|^2."""
    # one rating r=3 at (0,0); p=(1,0.5), q=(1,2); dot = 2, e = 1, lr = 0.1, lp=lq=0.1
    P = np.array([[1.0, 0.5]], np.float32)
    Q = np.array([[1.0, 2.0]], np.float32)
    lr = np.full(4, 0.1, np.float32)
    P1, Q1, _, _ = oracle.sgd_serial([0], [0], [3.0], P, Q, lr, 0, 1, k=2, G=1, prec=0, lambda_p=0.1, lambda_q=0.1,
                                     seed=0, batch_len=1)
    # p' = p + 0.1*(1*q - 0.1 p) = (1 + 0.1*(1-0.1), 0.5 + 0.1*(2-0.05)) = (1.09, 0.695)
    # q' = q + 0.1*(1*p - 0.1 q) = (1 + 0.1*(1-0.1), 2 + 0.1*(0.5-0.2)) = (1.09, 2.03)
    assert np.allclose(P1, [[1.09, 0.695]], atol=1e-6) and np.allclose(Q1, [[1.09, 2.03]], atol=1e-6)
    # gradient check at a random point with k=8
    rng = np.random.default_rng(8)
    p = rng.normal(size=(1, 8)).astype(np.float64)
    q = rng.normal(size=(1, 8)).astype(np.float64)
    r, lp, lq, eta = 1.7, 0.05, 0.02, 1e-3

    def loss(pp, qq):
        e = r - pp @ qq
        return 0.5 * e * e + 0.5 * lp * pp @ pp + 0.5 * lq * qq @ qq

    g_p = np.array([(loss(p[0] + h, q[0]) - loss(p[0] - h, q[0])) / 2e-6 for h in np.eye(8) * 1e-6])
    g_q = np.array([(loss(p[0], q[0] + h) - loss(p[0], q[0] - h)) / 2e-6 for h in np.eye(8) * 1e-6])
    P2, Q2 = oracle.sgd_serial_f64([0], [0], [r], p, q, np.full(2, eta, np.float32), 0, 1, k=8, lambda_p=lp,
                                   lambda_q=lq, seed=0, batch_len=1)
    lr32 = float(np.float32(eta))
    assert np.allclose((P2[0] - p[0]) / -lr32, g_p, rtol=1e-5, atol=1e-7)
    assert np.allclose((Q2[0] - q[0]) / -lr32, g_q, rtol=1e-5, atol=1e-7)
    P3, Q3, _, _ = oracle.sgd_serial([0], [0], [r], p.astype(np.float32), q.astype(np.float32),
                                     np.full(2, eta, np.float32), 0, 1, k=8, G=8, prec=0, lambda_p=lp, lambda_q=lq,
                                     seed=0, batch_len=1)
    assert np.allclose(P3, P2, atol=1e-6) and np.allclose(Q3, Q2, atol=1e-6)


def test_serial_sgd_fits_a_rank_one_problem():
    rng = np.random.default_rng(9)
    m, n, k = 60, 40, 8
    a = rng.uniform(0.5, 1.5, m)
    b = rng.uniform(0.5, 1.5, n)
    uu, vv = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    u, v = uu.ravel().astype(np.int32), vv.ravel().astype(np.int32)
    r = (a[u] * b[v]).astype(np.float32)
    P, _ = oracle.init_factors(m, k, 0, 3, 0)
    Q, _ = oracle.init_factors(n, k, 0, 3, 1)
    lr = np.full(60, 0.05, np.float32)
    P1, Q1, _, _ = oracle.sgd_serial(u, v, r, P, Q, lr, 0, 60, k=k, G=4, prec=0, lambda_p=0.0, lambda_q=0.0, seed=3,
                                     batch_len=16)
    fit = np.sqrt(np.mean((np.sum(P1[u] * Q1[v], 1) - r) ** 2))
    assert fit < 0.01


def test_serial_windows_follow_the_window_draws():
    """With lr = 0 an epoch changes nothing (the decay 1 - lr*lambda is then
    exactly 1), and the window starts of an epoch differ.  Which records a
    window visits is brute-forced in
    test_serial_windows_visit_L_consecutive_records_modulo_nnz."""
    rng = np.random.default_rng(10)
    m, n, k, nnz = 20, 10, 4, 50
    u = rng.integers(0, m, nnz).astype(np.int32)
    v = rng.integers(0, n, nnz).astype(np.int32)
    r = np.ones(nnz, np.float32)
    P, _ = oracle.init_factors(m, k, 0, 1, 0)
    Q, _ = oracle.init_factors(n, k, 0, 1, 1)
    P1, Q1, _, _ = oracle.sgd_serial(u, v, r, P, Q, np.zeros(3, np.float32), 0, 2, k=k, G=2, prec=0, lambda_p=0.1,
                                     lambda_q=0.1, seed=5, batch_len=7)
    assert np.array_equal(P1, P) and np.array_equal(Q1, Q)
    starts = [oracle.window_start(5, 0, 0, 0, 0, b, nnz) for b in range(math.ceil(nnz / 7))]
    assert len(set(starts)) > 1


@pytest.mark.parametrize("prec", [0, 1, 4])
def test_block_windows_restate_the_serial_epoch(prec):
    """One rating block's windows (the stratum schedule's unit) against the
    serialized epoch: with tag 0 the windows [0, ceil(nnz/L)) are that epoch
    bit for bit (same draws, same rounding words), split anywhere into two
    calls they are too, and another block id draws other windows."""
    rng = np.random.default_rng(11)
    m, n, k, nnz, L = 40, 30, 16, 700, 32
    u = rng.integers(0, m, nnz).astype(np.int32)
    v = rng.integers(0, n, nnz).astype(np.int32)
    r = rng.uniform(1, 5, nnz).astype(np.float32)
    P0, _ = oracle.init_factors(m, k, prec, 3, 0)
    Q0, _ = oracle.init_factors(n, k, prec, 3, 1)
    lr = oracle.lr_table(0.08, 0.3, 0.0, 4)
    kw = dict(k=k, G=4, prec=prec, lambda_p=0.05, lambda_q=0.04, seed=3, batch_len=L)
    nb = math.ceil(nnz / L)
    Ps, Qs, _, _ = oracle.sgd_serial(u, v, r, P0, Q0, lr, 1, 2, **kw)
    P, Q = P0.copy(), Q0.copy()
    oracle.sgd_block_windows(u, v, r, P, Q, lr, 1, 0, 0, nb, **kw)
    assert np.array_equal(P, Ps) and np.array_equal(Q, Qs)
    P2, Q2 = P0.copy(), Q0.copy()
    oracle.sgd_block_windows(u, v, r, P2, Q2, lr, 1, 0, 0, 9, **kw)
    oracle.sgd_block_windows(u, v, r, P2, Q2, lr, 1, 0, 9, nb - 9, **kw)
    assert np.array_equal(P2, Ps) and np.array_equal(Q2, Qs)
    P3, Q3 = P0.copy(), Q0.copy()
    oracle.sgd_block_windows(u, v, r, P3, Q3, lr, 1, 5, 0, nb, **kw)
    assert not np.array_equal(P3, Ps)


# ---------------------------------------------------------------- evaluation vs the reference binary
@pytest.mark.skipif(not oracle.have_mf_predict(), reason="reference mf-predict not built")
def test_rmse_and_model_file_match_reference_mf_predict(tmp_path):
    from cumf_sgd_amd import datagen

    rng = np.random.default_rng(11)
    m, n, k = 50, 30, 16
    P = rng.normal(0, 0.5, (m, k)).astype(np.float32)
    Q = rng.normal(0, 0.5, (n, k)).astype(np.float32)
    tu = rng.integers(-2, m + 3, 3000).astype(np.int32)
    tv = rng.integers(-2, n + 3, 3000).astype(np.int32)
    tr = rng.integers(1, 6, 3000).astype(np.float32)
    b = np.float32(3.3)
    model = str(tmp_path / "model")
    test = str(tmp_path / "test.bin")
    assert oracle.save_model(model, P, Q, b) == 0
    datagen.write_triplets(test, tu, tv, tr)
    ref = oracle.mf_predict_rmse(test, model)
    # mf-predict drops negative ids? it reads them as-is and predicts b (u < 0)
    mine = oracle.rmse(tu, tv, tr, P, Q, b)
    assert ref == f"{mine:.4f}"
    # byte layout: 20-byte header, then P rows, then Q rows
    raw = open(model, "rb").read()
    assert len(raw) == 20 + 4 * (m + n) * k
    assert np.array_equal(np.frombuffer(raw[:16], "<i4"), [0, m, n, k])
    assert np.frombuffer(raw[16:20], "<f4")[0] == b


def _fmt4(x):
    return "nan" if np.isnan(x) else f"{x:.4f}"


def test_evaluation_criteria_match_committed_mf_predict_vectors():
    """RMSE, MAE, GKL, LOGLOSS and ACC of the oracle against the 4-decimal
    values the reference's compiled mf-predict printed for the committed
    cases (tests/golden/make_mf_predict_vectors.py): ratings 1-5 with
    out-of-range ids, +-1 labels with a NaN factor row, a wide value range,
    and row/column MPR and AUC on distinct +-1 pairs past the model's extent."""
    g = json.load(open(os.path.join(GOLDEN, "mf_predict_vectors.json")))
    assert len(g["cases"]) == 4 and "row_auc" in g["cases"][3]
    for case in g["cases"]:
        d = np.load(os.path.join(GOLDEN, case["file"]))
        for key, crit in (("rmse", 0), ("mae", 1), ("gkl", 2), ("logloss", 5), ("acc", 6), ("row_mpr", 10),
                          ("col_mpr", 11), ("row_auc", 12), ("col_auc", 13)):
            if key not in case:
                continue
            mine = oracle.evaluate(crit, d["tu"], d["tv"], d["tr"], d["P"], d["Q"], np.float32(d["b"]))
            want = case[key].replace("-nan", "nan")
            assert _fmt4(mine) == want, (case["file"], key, mine, want)


@pytest.mark.skipif(not oracle.have_mf_predict(), reason="reference mf-predict not built")
def test_evaluation_criteria_match_reference_mf_predict(tmp_path):
    """Fresh random models against the live compiled evaluator, every -e."""
    from cumf_sgd_amd import datagen

    rng = np.random.default_rng(17)
    for trial in range(3):
        m, n, k = 40 + trial, 25 + trial, (8, 16, 32)[trial]
        P = np.abs(rng.normal(0.5, 0.3, (m, k))).astype(np.float32)
        Q = np.abs(rng.normal(0.5, 0.3, (n, k))).astype(np.float32)
        tu = rng.integers(-1, m + 2, 2500).astype(np.int32)
        tv = rng.integers(0, n, 2500).astype(np.int32)
        tr = rng.integers(1, 6, 2500).astype(np.float32)
        model, test = str(tmp_path / "m"), str(tmp_path / "t.bin")
        oracle.save_model(model, P, Q, np.float32(2.5))
        datagen.write_triplets(test, tu, tv, tr)
        for crit in (0, 1, 2, 5, 6):
            assert oracle.mf_predict_eval(test, model, crit) == _fmt4(
                oracle.evaluate(crit, tu, tv, tr, P, Q, np.float32(2.5))), (trial, crit)


# ---------------------------------------------------------------- window semantics, remap, rescale, model scale
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6, 7])
def test_serial_windows_visit_L_consecutive_records_modulo_nnz(L):
    """Brute force of the window semantics (singleGPU/sgd_k128_kernel_hogwild_warp32.h:41-56,
    worker count 1, update_count = ceil(nnz/L), sgd_kernel.cu:197-208): every
    record has its own user and item, q = 0, r = 0, lambda = 1 and lr = 0.5,
    so an update only multiplies the record's p by 1 - lr*lambda = 1/2 and
    p ends at exactly 2^-(number of updates of that record).  The counts
    must equal those of the drawn windows: window b of epoch ep covers the
    records (start_b + i) mod nnz, i = 0..L-1."""
    nnz, k, seed, epochs = 23, 2, 7 + L, 2
    u = np.arange(nnz, dtype=np.int32)
    v = np.arange(nnz, dtype=np.int32)
    r = np.zeros(nnz, np.float32)
    P0 = np.ones((nnz, k), np.float32)
    Q0 = np.zeros((nnz, k), np.float32)
    lr = np.full(epochs, 0.5, np.float32)
    P1, Q1, _, _ = oracle.sgd_serial(u, v, r, P0, Q0, lr, 0, epochs, k=k, G=2, prec=oracle.F32, lambda_p=1.0,
                                     lambda_q=1.0, seed=seed, batch_len=L)
    assert np.array_equal(Q1, Q0)
    assert np.all(P1[:, 0] == P1[:, 1])
    got = -np.log2(P1[:, 0].astype(np.float64))
    assert np.array_equal(got, np.round(got))
    want = np.zeros(nnz, np.int64)
    for ep in range(epochs):
        for b in range(-(-nnz // L)):
            s = oracle.window_start(seed, 0, 0, 0, ep, b, nnz)
            for i in range(L):
                want[(s + i) % nnz] += 1
    assert want.sum() == epochs * L * -(-nnz // L)
    assert np.array_equal(got.astype(np.int64), want)


def test_serial_windows_walk_records_in_order():
    """Within a window the records are applied in index order: two records of
    the same user and item, r = 1 then r = 3 (lambda 0, lr 1/2, k = 1 lane
    group of 2 with q = (1, 0), p = (0, 0)): the first update gives
    p0 = 1/2, q unchanged (p was 0); the second sees the dot 1/2, e = 5/2
    and moves p0 by 5/4 and q0 by 5/8.  The reverse order ends elsewhere."""
    u = np.zeros(2, np.int32)
    v = np.zeros(2, np.int32)
    P0 = np.zeros((1, 2), np.float32)
    Q0 = np.array([[1.0, 0.0]], np.float32)
    lr = np.full(1, 0.5, np.float32)
    kw = dict(k=2, G=2, prec=oracle.F32, lambda_p=0.0, lambda_q=0.0, batch_len=2)
    # pick a seed whose single window starts at record 0 (order 0, 1)
    seed = next(s for s in range(1000) if oracle.window_start(s, 0, 0, 0, 0, 0, 2) == 0)
    P1, Q1, _, _ = oracle.sgd_serial(u, v, np.float32([1, 3]), P0, Q0, lr, 0, 1, seed=seed, **kw)
    assert P1[0].tolist() == [0.5 + 1.25, 0.0] and Q1[0].tolist() == [1.0 + 0.625, 0.0]
    P2, Q2, _, _ = oracle.sgd_serial(u, v, np.float32([3, 1]), P0, Q0, lr, 0, 1, seed=seed, **kw)
    # r = 3 first: p0 = 3/2, then dot 3/2, e = -1/2: p0 = 3/2 - 1/4, q0 = 1 - 3/8
    assert P2[0].tolist() == [1.25, 0.0] and Q2[0].tolist() == [0.625, 0.0]


def test_shuffle_problem_is_the_gather_by_the_maps():
    """shuffle_problem (singleGPU/sgd.cu:126-134) against numpy's gather
    u <- p_map[u], v <- q_map[v], values untouched."""
    rng = np.random.default_rng(21)
    m, n, nnz = 1000, 300, 20_000
    nodes = np.zeros(nnz, oracle.NODE_DTYPE)
    nodes["u"] = rng.integers(0, m, nnz)
    nodes["v"] = rng.integers(0, n, nnz)
    nodes["rate"] = rng.uniform(1, 5, nnz).astype(np.float32)
    pm = rng.permutation(m).astype(np.int32)
    qm = rng.permutation(n).astype(np.int32)
    out = oracle.shuffle_problem(nodes, pm, qm)
    assert np.array_equal(out["u"], pm[nodes["u"]])
    assert np.array_equal(out["v"], qm[nodes["v"]])
    assert np.array_equal(out["rate"].view(np.uint32), nodes["rate"].view(np.uint32))


def test_scale_model_multiplies_by_float_sqrt_of_scale():
    """scale_model (singleGPU/sgd.cu:593-606): every element times
    float(sqrt(scale)), against numpy's correctly rounded float32 sqrt."""
    rng = np.random.default_rng(22)
    vec = rng.normal(0, 1, (37, 16)).astype(np.float32)
    for scale in (1.0, 1.08918, 34.6, 1e-4, 2.0):
        fs = np.sqrt(np.float32(scale))
        assert np.array_equal(oracle.scale_model(vec, scale), vec * fs)


def _e4m3_values():
    torch = pytest.importorskip("torch")
    codes = np.arange(256, dtype=np.uint8)
    vals = torch.from_numpy(codes).view(torch.float8_e4m3fn).float().numpy()
    return torch, codes, vals


def test_fp8_rescale_against_torch_e4m3():
    """The between-epoch FP8 rescale against facts outside the oracle: the new
    power-of-two scale puts the row's largest magnitude in (28, 56]; the new
    bytes are torch's round-to-nearest-even float8_e4m3fn of old value / new
    scale; when the scale does not grow every value is kept exactly (a power
    of two times an E4M3 value below 56 is an E4M3 value); when it grows each
    value moves by at most half an E4M3 quantum at the new scale; all-zero
    rows are left alone."""
    torch, codes, vals = _e4m3_values()
    rng = np.random.default_rng(23)
    rows, k = 400, 32
    finite = codes[np.isfinite(vals)]
    data = rng.choice(finite, (rows, k)).astype(np.uint8)
    # rows of small magnitude too (scale must grow or shrink by several octaves)
    small = np.abs(vals[finite]) <= 2
    data[::3] = rng.choice(finite[small], (len(range(0, rows, 3)), k))
    data[5] = 0                                  # an all-zero row
    scales = np.ldexp(np.float32(1), rng.integers(-20, 8, rows)).astype(np.float32)
    old = vals[data] * scales[:, None]
    new_d, new_s = oracle.fp8_rescale(data, scales)
    assert new_s[5] == scales[5] and np.array_equal(new_d[5], data[5])
    live = np.abs(old).max(1) > 0
    amax = np.abs(old).max(1)
    assert np.all(np.log2(new_s[live]) == np.round(np.log2(new_s[live])))
    assert np.all(amax[live] / new_s[live] <= 56) and np.all(amax[live] / new_s[live] > 28)
    want = torch.from_numpy((old / new_s[:, None]).astype(np.float32)).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(new_d[live], want[live])
    got = vals[new_d] * new_s[:, None]
    shrink = live & (new_s <= scales)
    assert np.array_equal(got[shrink], old[shrink])
    grow = live & (new_s > scales)
    assert grow.sum() > 50 and shrink.sum() > 50
    # half a quantum at the new scale: E4M3 has 3 mantissa bits, subnormal
    # quantum 2^-9
    x = np.abs(old[grow] / new_s[grow, None])
    e = np.maximum(np.floor(np.log2(np.where(x > 0, x, 1))), -6)
    quantum = np.ldexp(1.0, (e - 3).astype(int)) * new_s[grow, None]
    assert np.all(np.abs(got[grow] - old[grow]) <= quantum / 2)
