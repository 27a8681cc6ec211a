"""cumf_train_noalloc (training without any device allocation by the library)
against cumf_train on a context holding the same ratings and factors: the
serialized schedule is deterministic, so the factors must agree bit for bit,
on two calls in a row whose second input is smaller (the fixed device array
keeps the first call's task record and learning-rate table, which the second
call must replace)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _to_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to("cuda")


def _from_dev(t, dtype):
    return t.cpu().numpy().view(dtype)


@pytest.mark.parametrize("prec", [0, 1, 3, 4])
def test_noalloc_equals_context_train_on_two_calls(gpu, prec):
    import torch

    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    k = 128
    dt = cs.STORAGE_DTYPE[prec]
    stream = torch.cuda.current_stream().cuda_stream
    # the second call is smaller in every dimension and starts at a later epoch
    for call, (m, n, nnz, e0, e1) in enumerate(((700, 260, 40_000, 0, 3), (310, 90, 6_000, 3, 4))):
        print(f"[noalloc] prec={prec} call={call} m={m} n={n} nnz={nnz} epochs=[{e0},{e1})", flush=True)
        R = datagen.uniform(50 + call, m, n, nnz)
        u, v, r = R.numpy()[:3]
        opts = cs.make_opts(k=k, schedule=cs.SERIAL, seed=11 + call, lambda_p=0.04, lambda_q=0.05)
        with cs.Context(0) as c:
            c.set_ratings(u, v, r, m, n)
            c.init_factors(k, prec, seed=7 + call)
            P0, ps0 = c.get_factors(0)
            Q0, qs0 = c.get_factors(1)
            c.train(opts, e0, e1)
            P1, ps1 = c.get_factors(0)
            Q1, qs1 = c.get_factors(1)
        du = torch.from_numpy(u).to("cuda")
        dv = torch.from_numpy(v).to("cuda")
        dr = torch.from_numpy(r).to("cuda")
        dP, dQ = _to_dev(P0), _to_dev(Q0)
        dps = torch.from_numpy(ps0).to("cuda") if prec == cs.FP8_E4M3 else None
        dqs = torch.from_numpy(qs0).to("cuda") if prec == cs.FP8_E4M3 else None
        torch.cuda.synchronize()
        cs.train_noalloc(opts, prec, du, dv, dr, m, n, dP, dQ, e0, e1, ps=dps, qs=dqs, stream=stream)
        torch.cuda.synchronize()
        assert np.array_equal(_from_dev(dP, dt).reshape(m, k).view(np.uint8), P1.view(np.uint8))
        assert np.array_equal(_from_dev(dQ, dt).reshape(n, k).view(np.uint8), Q1.view(np.uint8))
        if prec == cs.FP8_E4M3:
            assert np.array_equal(dps.cpu().numpy(), ps1) and np.array_equal(dqs.cpu().numpy(), qs1)
        assert not np.array_equal(P1.view(np.uint8), P0.view(np.uint8))


def test_noalloc_hogwild_trains_and_rejects_what_it_cannot_hold(gpu):
    import torch

    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    k, m, n = 64, 500, 200
    R = datagen.uniform(60, m, n, 50_000, 5_000)
    u, v, r, tu, tv, tr = R.numpy()
    with cs.Context(0) as c:
        c.set_ratings(u, v, r, m, n)
        c.set_test(tu, tv, tr)
        c.init_factors(k, cs.FP16, seed=2)
        P0, _ = c.get_factors(0)
        Q0, _ = c.get_factors(1)
        before = c.test_rmse(1.0, 3.0)
        dP, dQ = _to_dev(P0), _to_dev(Q0)
        du, dv, dr = (torch.from_numpy(x).to("cuda") for x in (u, v, r))
        opts = cs.make_opts(k=k, schedule=cs.HOGWILD, seed=4)
        print("[noalloc] hogwild fp16 k=64, 5 epochs", flush=True)
        cs.train_noalloc(opts, cs.FP16, du, dv, dr, m, n, dP, dQ, 0, 5)
        torch.cuda.synchronize()
        c.set_factors(0, _from_dev(dP, np.uint16).reshape(m, k))
        c.set_factors(1, _from_dev(dQ, np.uint16).reshape(n, k))
        after = c.test_rmse(1.0, 3.0)
        assert np.isfinite(after) and after < before - 0.05
        # sizes and schedules the fixed array cannot hold are refused before any launch
        with pytest.raises(cs.CumfError, match="beyond the allocation-free limit"):
            cs.train_noalloc(opts, cs.FP16, du, dv, dr, m, n, dP, dQ, 0, cs.NOALLOC_MAX_EPOCHS + 1)
        with pytest.raises(cs.CumfError, match="HOGWILD or SERIAL"):
            cs.train_noalloc(cs.make_opts(k=k, schedule=cs.STRATUM), cs.FP16, du, dv, dr, m, n, dP, dQ, 0, 1)
        # the largest epoch count is accepted
        cs.train_noalloc(opts, cs.FP16, du, dv, dr, m, n, dP, dQ, cs.NOALLOC_MAX_EPOCHS - 1,
                         cs.NOALLOC_MAX_EPOCHS)
        torch.cuda.synchronize()
        assert np.all(np.isfinite(_from_dev(dQ, np.float16)))
