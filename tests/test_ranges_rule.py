"""The id-range cut of cumf_partition_strata for s <= 8, as restated in
ranges_rule.py, against brute force: over every contiguous cut of small
degree histograms its largest range is the least possible, and so is the
capacity its binary search finds.  (The GPU test compares the library's
cut with the restatement.)"""
import itertools

import numpy as np

from ranges_rule import capped_nearest_ranges, least_largest_range


def _optimum(deg, s):
    """Least largest range over all cuts of deg into s contiguous ranges
    (empty ranges allowed)."""
    c = np.concatenate([[0], np.cumsum(deg)])
    best = None
    for cuts in itertools.combinations_with_replacement(range(len(deg) + 1), s - 1):
        b = (0,) + cuts + (len(deg),)
        most = max(int(c[b[j + 1]] - c[b[j]]) for j in range(s))
        best = most if best is None else min(best, most)
    return best


def test_capped_nearest_cut_reaches_the_least_largest_range():
    rng = np.random.default_rng(7)
    weights = np.array([0, 1, 1, 2, 3, 5, 8, 13, 30])
    for _ in range(3000):
        n = int(rng.integers(1, 10))
        s = int(rng.integers(2, 7))
        deg = rng.choice(weights, n)
        if deg.sum() == 0:
            continue
        sizes = capped_nearest_ranges(deg, s)
        opt = _optimum(deg, s)
        assert len(sizes) == s and sizes.sum() == deg.sum() and np.all(sizes >= 0)
        assert least_largest_range(deg, s) == opt, (deg, s)
        assert sizes.max() == opt, (deg, s, sizes, opt)


def test_nearer_share_alone_would_pass_the_capacity():
    """Why the cut checks that the rest still fits: [1, 3, 1, 0, 5, 1] in 3
    ranges at M = 5.  Closing at every boundary nearer the even share closes
    [1, 3] (4, nearer 11/3 than 5 is), then [1, 0] before the 5 (which would
    pass M), and leaves [5, 1] = 6.  With the check, [1, 3] stays open
    (the rest, [1, 0, 5, 1], would take 3 ranges of at most 5 with 2 left)
    and the cut is [1, 3, 1] [0, 5] [1]."""
    deg = np.array([1, 3, 1, 0, 5, 1])

    def unchecked(deg, s, M):
        sizes, run, left = [], 0, int(deg.sum())
        for d in deg.tolist():
            k = s - len(sizes)
            if k > 1 and run > 0 and (run + d > M or 2 * run * k + d * k > 2 * left):
                sizes.append(run)
                left -= run
                run = 0
            run += d
        return sizes + [run]

    assert least_largest_range(deg, 3) == _optimum(deg, 3) == 5
    assert unchecked(deg, 3, 5) == [4, 1, 6]
    assert capped_nearest_ranges(deg, 3).tolist() == [5, 5, 1]
