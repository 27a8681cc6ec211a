"""Restatement of cumf_partition_strata's id-range cuts from a degree
histogram (cumf_sgd_amd/csrc/capi.hip, balanced_bounds), shared by the CPU
test that pins the rule by brute force and the GPU test that compares the
library's cut with it.  No counterpart in the reference, whose grid_problem
cuts equal id segments (singleGPU/sgd.cu:419-423)."""
import numpy as np


def first_crossing_ranges(deg, s):
    """Range sizes when range j ends at the first row whose running count
    reaches j/s of the ratings (the cut for s > 8)."""
    nnz = int(deg.sum())
    b = [0]
    c = np.cumsum(deg)
    for j in range(1, s):
        hit = np.flatnonzero(c * s >= nnz * j)
        b.append(int(hit[0]) + 1 if len(hit) else len(deg))
    b.append(len(deg))
    b = np.maximum.accumulate(np.array(b))
    return np.diff(np.concatenate([[0], c])[b])


def fill_count(deg, M):
    """Ranges of at most M that rows filled in id order take, each closed
    before the row that would take it past M."""
    n, run = 1, 0
    for d in deg:
        if run + d > M:
            n, run = n + 1, 0
        run += d
    return n


def least_largest_range(deg, s):
    """The least count M for which fill_count(deg, M) <= s."""
    lo, hi = max(int(deg.max()), -(-int(deg.sum()) // s)), max(int(deg.sum()), 1)
    while lo < hi:
        mid = (lo + hi) // 2
        if fill_count(deg, mid) <= s:
            hi = mid
        else:
            lo = mid + 1
    return lo


def capped_nearest_ranges(deg, s):
    """Range sizes of the cut for s <= 8: a range closes before the row that
    would take it past M = least_largest_range, or at the row boundary
    nearer the remaining ratings' even share when the rows from there still
    fill at most the ranges left at M."""
    deg = [int(d) for d in deg]
    M = least_largest_range(np.asarray(deg), s)
    sizes, run, left = [], 0, sum(deg)
    for i, d in enumerate(deg):
        k = s - len(sizes)
        if k > 1 and run > 0:
            nearer = 2 * run * k + d * k > 2 * left
            if run + d > M or (nearer and fill_count(deg[i:], M) <= k - 1):
                sizes.append(run)
                left -= run
                run = 0
        run += d
    sizes.append(run)
    return np.array(sizes + [0] * (s - len(sizes)))
