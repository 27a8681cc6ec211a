"""Seeded synthetic rating matrices for the benchmark and the tests.

No real data is available, so every workload is generated: ids, degrees and
values follow the shapes of the reference's target datasets (Netflix-like and
larger), with a low-rank ground truth so that the test RMSE is learnable.
Generation runs in torch on the given device (the GPU for the large shapes).
Records come out in random order, as three arrays (u int32, v int32, r float32).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class Ratings:
    m: int
    n: int
    u: torch.Tensor
    v: torch.Tensor
    r: torch.Tensor
    tu: torch.Tensor
    tv: torch.Tensor
    tr: torch.Tensor

    @property
    def nnz(self) -> int:
        return int(self.u.numel())

    def numpy(self):
        f = lambda t: t.detach().cpu().numpy()
        return f(self.u), f(self.v), f(self.r), f(self.tu), f(self.tv), f(self.tr)


def _zipf_sample(gen, n_items: int, s: float, count: int, device) -> torch.Tensor:
    """Ranks in [0, n_items) with P(rank i) proportional to (i+1)^-s."""
    w = torch.arange(1, n_items + 1, dtype=torch.float64, device=device).pow(-s)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    x = torch.rand(count, generator=gen, dtype=torch.float64, device=device)
    return torch.searchsorted(cdf, x).clamp_(max=n_items - 1)


def _scores(gen, u, v, m, n, rank, device, chunk=1 << 24) -> torch.Tensor:
    U = torch.randn(m, rank, generator=gen, device=device)
    V = torch.randn(n, rank, generator=gen, device=device)
    out = torch.empty(u.numel(), dtype=torch.float32, device=device)
    for s in range(0, u.numel(), chunk):
        e = min(s + chunk, u.numel())
        out[s:e] = (U[u[s:e].long()] * V[v[s:e].long()]).sum(1)
    return out


def _split(gen, u, v, r, n_test, device) -> tuple:
    perm = torch.randperm(u.numel(), generator=gen, device=device)
    u, v, r = u[perm], v[perm], r[perm]
    return (u[n_test:].contiguous(), v[n_test:].contiguous(), r[n_test:].contiguous(),
            u[:n_test].contiguous(), v[:n_test].contiguous(), r[:n_test].contiguous())


def small(seed: int = 1, device="cpu", m: int = 10_000, n: int = 2_000, nnz: int = 1_000_000,
          test_frac: float = 0.1) -> Ratings:
    """Plain method, small: 10,000 x 2,000, 1,000,000 distinct ratings.

    Row and column ranks are Zipf(1.1) over random id orders; duplicates are
    rejected.  r = clip(round(3.6 + U_u.V_v/4 + N(0, 0.5^2)), 1, 5) with rank-16
    Gaussian U, V; 90/10 random split.
    """
    gen = torch.Generator(device=device).manual_seed(seed)
    pu = torch.randperm(m, generator=gen, device=device)
    pv = torch.randperm(n, generator=gen, device=device)
    keys = torch.empty(0, dtype=torch.int64, device=device)
    draw = nnz
    while keys.numel() < nnz:
        ru = pu[_zipf_sample(gen, m, 1.1, draw, device)]
        rv = pv[_zipf_sample(gen, n, 1.1, draw, device)]
        new = ru.long() * n + rv.long()
        allk = torch.cat([keys, new])
        uniq, inv = torch.unique(allk, return_inverse=True)
        # keep first occurrences in draw order
        first = torch.full((uniq.numel(),), allk.numel(), dtype=torch.int64, device=device)
        first.scatter_reduce_(0, inv, torch.arange(allk.numel(), device=device), reduce="amin")
        keys = allk[torch.sort(first).values]
        draw *= 2
    keys = keys[:nnz]
    u = (keys // n).to(torch.int32)
    v = (keys % n).to(torch.int32)
    s = _scores(gen, u, v, m, n, 16, device) / 4.0
    noise = torch.randn(nnz, generator=gen, device=device) * 0.5
    r = torch.clamp(torch.round(3.6 + s + noise), 1, 5).to(torch.float32)
    tu_, tv_, tr_, eu, ev, er = _split(gen, u, v, r, int(round(nnz * test_frac)), device)
    return Ratings(m, n, tu_, tv_, tr_, eu, ev, er)


def _lognormal_degrees(gen, m: int, n: int, total: int, median: float, sigma: float, device) -> torch.Tensor:
    z = torch.randn(m, generator=gen, dtype=torch.float64, device=device)
    d = torch.exp(np.log(median) + sigma * z).clamp_(1, n)
    d = torch.floor(d * (total / d.sum())).clamp_(1, n).long()
    diff = total - int(d.sum())
    # spread the remainder one rating at a time over random users with room
    while diff != 0:
        idx = torch.randint(0, m, (abs(diff),), generator=gen, device=device)
        step = 1 if diff > 0 else -1
        room = (d[idx] < n) if step > 0 else (d[idx] > 1)
        idx = torch.unique(idx[room])
        idx = idx[: abs(diff)]
        d[idx] += step
        diff -= step * idx.numel()
    return d


def medium(seed: int = 2, device="cuda", m: int = 480_189, n: int = 17_770, n_train: int = 99_072_112,
           n_test: int = 1_408_395) -> Ratings:
    """Medium dense-ish (Netflix shape): 480,189 x 17,770, 99,072,112 train +
    1,408,395 test ratings.

    User degrees log-normal (median 100, sigma 1.215) scaled to the total and
    capped at n; items Zipf(0.9) over a random id order; values from a rank-16
    latent score plus N(0, 0.5^2) noise, mapped by quantiles onto 1..5 with
    P = .05/.10/.29/.34/.22; a uniformly random test split of the given size.
    Pairs may repeat (no de-duplication at this size).
    """
    total = n_train + n_test
    gen = torch.Generator(device=device).manual_seed(seed)
    deg = _lognormal_degrees(gen, m, n, total, 100.0, 1.215, device)
    u = torch.repeat_interleave(torch.arange(m, dtype=torch.int32, device=device), deg)
    pv = torch.randperm(n, generator=gen, device=device).to(torch.int32)
    v = pv[_zipf_sample(gen, n, 0.9, total, device)]
    s = _scores(gen, u, v, m, n, 16, device) / 4.0
    s += torch.randn(total, generator=gen, device=device) * 0.5
    probe = s[torch.randint(0, total, (1 << 20,), generator=gen, device=device)]
    cuts = torch.quantile(probe.double(), torch.tensor([0.05, 0.15, 0.44, 0.78], dtype=torch.float64, device=device))
    r = (torch.bucketize(s.double(), cuts) + 1).to(torch.float32)
    del s
    tu_, tv_, tr_, eu, ev, er = _split(gen, u, v, r, n_test, device)
    return Ratings(m, n, tu_, tv_, tr_, eu, ev, er)


def uniform(seed: int, m: int, n: int, nnz: int, n_test: int = 0, device="cpu") -> Ratings:
    """Tiny uniform problems for unit tests: integer ratings 1..5 from a rank-4
    score, ids uniform."""
    gen = torch.Generator(device=device).manual_seed(seed)
    u = torch.randint(0, m, (nnz + n_test,), generator=gen, device=device, dtype=torch.int32)
    v = torch.randint(0, n, (nnz + n_test,), generator=gen, device=device, dtype=torch.int32)
    s = _scores(gen, u, v, m, n, 4, device) / 2.0
    r = torch.clamp(torch.round(3.0 + s), 1, 5).to(torch.float32)
    tu_, tv_, tr_, eu, ev, er = _split(gen, u, v, r, n_test, device)
    return Ratings(m, n, tu_, tv_, tr_, eu, ev, er)


def write_triplets(path: str, u, v, r) -> None:
    """Write the reference's binary input: (int32 u, int32 v, float32 r) records."""
    rec = np.empty(len(u), dtype=[("u", "<i4"), ("v", "<i4"), ("r", "<f4")])
    rec["u"], rec["v"], rec["r"] = u, v, r
    rec.tofile(path)


@dataclass
class Problem:
    m: int
    n: int
    k: int
    u: torch.Tensor
    v: torch.Tensor
    r: torch.Tensor
    tu: torch.Tensor
    tv: torch.Tensor
    tr: torch.Tensor


def batched(seed: int = 4, count: int = 256, device="cpu", users=(1_000, 200_000), items=(500, 50_000),
            density=(0.001, 0.02), ks=(32, 64, 128), test_frac: float = 0.1) -> list:
    """Batched: `count` independent problems, users / items / density each
    log-uniform over the given ranges, k drawn from `ks`, ratings 1..5 from a
    rank-16 score + N(0, 0.5^2) noise as in the small config, uniform ids,
    random 90/10 splits."""
    gen = torch.Generator(device="cpu").manual_seed(seed)

    def logu(lo, hi):
        return float(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * torch.rand(1, generator=gen).item()))

    out = []
    for i in range(count):
        m = int(round(logu(*users)))
        n = int(round(logu(*items)))
        d = logu(*density)
        k = int(ks[int(torch.randint(0, len(ks), (1,), generator=gen).item())])
        nnz = max(20, int(round(m * n * d)))
        g2 = torch.Generator(device=device).manual_seed(seed * 100_003 + i)
        u = torch.randint(0, m, (nnz,), generator=g2, device=device, dtype=torch.int32)
        v = torch.randint(0, n, (nnz,), generator=g2, device=device, dtype=torch.int32)
        s = _scores(g2, u, v, m, n, 16, device) / 4.0
        r = torch.clamp(torch.round(3.6 + s + torch.randn(nnz, generator=g2, device=device) * 0.5), 1, 5)
        tu_, tv_, tr_, eu, ev, er = _split(g2, u, v, r.to(torch.float32), int(round(nnz * test_frac)), device)
        out.append(Problem(m, n, k, tu_, tv_, tr_, eu, ev, er))
    return out


# 11 bin masses of the large sparse config's values 0, 10, ..., 100: a mode at
# 0 and a second mode at 90.
BIMODAL_MASSES = (0.16, 0.04, 0.03, 0.03, 0.04, 0.06, 0.08, 0.11, 0.15, 0.18, 0.12)


def large_sparse(seed: int = 3, device="cuda", m: int = 1_000_990, n: int = 624_961, nnz: int = 252_800_275,
                 test_frac: float = 0.01, chunk: int = 1 << 25) -> Ratings:
    """Large sparse: 1,000,990 x 624,961, 252,800,275 ratings, Zipf(1.0) ranks
    on both axes over random id orders, float32 values in {0, 10, ..., 100}
    mapped by quantiles of a rank-16 score + noise onto BIMODAL_MASSES, 1%
    random test split.  Pairs may repeat."""
    gen = torch.Generator(device=device).manual_seed(seed)
    pu = torch.randperm(m, generator=gen, device=device).to(torch.int32)
    pv = torch.randperm(n, generator=gen, device=device).to(torch.int32)
    u = torch.empty(nnz, dtype=torch.int32, device=device)
    v = torch.empty(nnz, dtype=torch.int32, device=device)
    for s in range(0, nnz, chunk):
        e = min(nnz, s + chunk)
        u[s:e] = pu[_zipf_sample(gen, m, 1.0, e - s, device)]
        v[s:e] = pv[_zipf_sample(gen, n, 1.0, e - s, device)]
    sc = _scores(gen, u, v, m, n, 16, device) / 4.0
    sc += torch.randn(nnz, generator=gen, device=device) * 0.5
    probe = sc[torch.randint(0, nnz, (1 << 20,), generator=gen, device=device)]
    cum = torch.tensor(np.cumsum(BIMODAL_MASSES)[:-1], dtype=torch.float64, device=device)
    cuts = torch.quantile(probe.double(), cum)
    r = torch.empty(nnz, dtype=torch.float32, device=device)
    for s in range(0, nnz, chunk):
        e = min(nnz, s + chunk)
        r[s:e] = torch.bucketize(sc[s:e].double(), cuts).to(torch.float32) * 10.0
    del sc
    tu_, tv_, tr_, eu, ev, er = _split(gen, u, v, r, int(round(nnz * test_frac)), device)
    return Ratings(m, n, tu_, tv_, tr_, eu, ev, er)


def tall_into(u, v, r, m: int, n: int, seed: int = 5, rank: int = 32, chunk: int = 1 << 26, test=None):
    """Fill device arrays u, v, r (length nnz) with the tall-skinny config:
    users uniform over m, items Zipf(1.1) over a random order of n, values
    exp(z) with z = U_u.V_v / sqrt(rank) (unit variance, U, V ~ N(0,1)^rank),
    i.e. log-normal(0, 1).  `test` = (tu, tv, tr) arrays are filled with
    further ratings of the same ground truth."""
    device = u.device
    gen = torch.Generator(device=device).manual_seed(seed)
    pv = torch.randperm(n, generator=gen, device=device).to(torch.int32)
    U = torch.randn(m, rank, generator=gen, device=device, dtype=torch.float32)
    V = torch.randn(n, rank, generator=gen, device=device, dtype=torch.float32)
    w = torch.arange(1, n + 1, dtype=torch.float64, device=device).pow(-1.1)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    inv = 1.0 / np.sqrt(rank)
    for (a, b, c) in ((u, v, r),) + ((tuple(test),) if test is not None else ()):
        nnz = a.numel()
        for s in range(0, nnz, chunk):
            e = min(nnz, s + chunk)
            uu = torch.randint(0, m, (e - s,), generator=gen, device=device, dtype=torch.int32)
            x = torch.rand(e - s, generator=gen, dtype=torch.float64, device=device)
            vv = pv[torch.searchsorted(cdf, x).clamp_(max=n - 1)]
            a[s:e] = uu
            b[s:e] = vv
            c[s:e] = torch.exp((U[uu.long()] * V[vv.long()]).sum(1) * inv)
    del U, V
