"""Error behaviour of the device path through the C ABI.

The reference has no status returns: `gpuErr` prints and calls exit
(singleGPU/sgd_kernel.h:6-14) and the reader exits on a missing file
(singleGPU/sgd.cu:224-234).  Here every entry point returns a code with the
last error's text, and every refusal happens on the host before a kernel is
launched, so a context stays usable after one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE = 1, 3


def _problem(m=300, n=200, nnz=20_000, n_test=2_000):
    from cumf_sgd_amd import datagen

    R = datagen.uniform(11, m, n, nnz, n_test)
    return R.numpy() + (m, n)


def _err(fn, code, text):
    import cumf_sgd_amd as cs

    with pytest.raises(cs.CumfError) as e:
        fn()
    assert e.value.code == code, str(e.value)
    assert text in str(e.value), str(e.value)


def test_out_of_order_calls_are_state_errors(gpu):
    import cumf_sgd_amd as cs

    u, v, r, tu, tv, tr, m, n = _problem()
    with cs.Context(0) as c:
        _err(lambda: c.init_factors(128, cs.FP16, 1), ERR_STATE, "set ratings")
        _err(lambda: c.partition_strata(4), ERR_STATE, "no ratings")
        _err(lambda: c.train(cs.make_opts(k=128), 0, 1), ERR_STATE, "no ratings")
        c.set_ratings(u, v, r, m, n)
        _err(lambda: c.train(cs.make_opts(k=128), 0, 1), ERR_STATE, "factors not initialised")
        c.init_factors(128, cs.FP16, 1)
        _err(lambda: c.test_rmse(), ERR_STATE, "no test ratings")
        # the stratum schedule needs an s x s partition first
        _err(lambda: c.train(cs.make_opts(k=128, schedule=cs.STRATUM), 0, 1), ERR_STATE, "s x s partition")
        c.set_test(tu, tv, tr)
        c.train(cs.make_opts(k=128, schedule=cs.HOGWILD, num_workers=64), 0, 2)
        assert np.isfinite(c.test_rmse(1.0, 3.0))


def test_bad_arguments_are_refused_and_the_context_stays_usable(gpu):
    import cumf_sgd_amd as cs

    u, v, r, tu, tv, tr, m, n = _problem()
    with cs.Context(0) as c:
        # a rating id outside [0, m) x [0, n): refused, and the ratings are
        # not usable until a valid set replaces them
        bad_u = u.copy()
        bad_u[123] = m
        _err(lambda: c.set_ratings(bad_u, v, r, m, n), ERR_ARG, "outside")
        _err(lambda: c.init_factors(128, cs.FP16, 1), ERR_STATE, "set ratings")
        bad_v = v.copy()
        bad_v[7] = -1
        _err(lambda: c.set_ratings(u, bad_v, r, m, n), ERR_ARG, "outside")

        c.set_ratings(u, v, r, m, n)
        c.set_test(tu, tv, tr)
        _err(lambda: c.init_factors(100, cs.FP16, 1), ERR_ARG, "k=100 unsupported")
        _err(lambda: c.init_factors(128, 9, 1), ERR_ARG, "bad precision")
        # a map that is not a permutation leaves the ratings as they were
        before = c.get_ratings()
        p_map = np.arange(m, dtype=np.int32)
        p_map[5] = p_map[6]
        _err(lambda: c.remap(p_map, np.arange(n, dtype=np.int32)), ERR_ARG, "not a permutation")
        after = c.get_ratings()
        assert all(np.array_equal(a, b) for a, b in zip(before, after))

        _err(lambda: c.partition_strata(0), ERR_ARG, "out of range")
        _err(lambda: c.partition_strata(65), ERR_ARG, "out of range")
        nnz = len(u)
        _err(lambda: c.set_blocks(2, 1, np.array([0, 5, nnz - 1], np.int64)), ERR_ARG, "span")
        _err(lambda: c.set_blocks(2, 1, np.array([0, nnz, 5], np.int64)), ERR_ARG, "span")
        _err(lambda: c.set_blocks(2, 1, np.array([0, nnz + 5, nnz], np.int64)), ERR_ARG, "ascending")
        _err(lambda: c.set_blocks(0, 1, np.array([0], np.int64)), ERR_ARG, "bad block grid")

        c.init_factors(64, cs.FP8_E4M3, 2)
        P, S = c.get_factors(0)
        _err(lambda: c.set_factors(0, P, None), ERR_ARG, "row scales")
        _err(lambda: c.train(cs.make_opts(k=128), 0, 1), ERR_ARG, "opts k=128")
        _err(lambda: c.train(cs.make_opts(k=64, pipeline_depth=3), 0, 1), ERR_ARG, "pipeline_depth")
        _err(lambda: c.train(cs.make_opts(k=64, batch_len=127), 0, 1), ERR_ARG, "even")
        _err(lambda: c.train(cs.make_opts(k=64, batch_len=4096), 0, 1), ERR_ARG, "out of range")
        _err(lambda: c.train(cs.make_opts(k=64, lanes_per_sample=3), 0, 1), ERR_ARG, "lanes_per_sample")
        _err(lambda: c.train(cs.make_opts(k=64, schedule=7), 0, 1), ERR_ARG, "bad schedule")
        _err(lambda: c.train(cs.make_opts(k=64, use_graph=2), 0, 1), ERR_ARG, "use_graph")
        _err(lambda: c.train(cs.make_opts(k=64), 3, 1), ERR_ARG, "epoch range")
        c.partition(2, 2)
        _err(lambda: c.train(cs.make_opts(k=64, schedule=cs.GRID), 0, 1, np.array([0, 1, 4, 2], np.int32)),
             ERR_ARG, "out of range")
        _err(lambda: c.train(cs.make_opts(k=64, schedule=cs.GRID), 0, 1, np.array([0, 1, 2], np.int32)),
             ERR_ARG, "block_order has 3 entries")
        _err(lambda: c.test_eval(3), ERR_ARG, "criterion 3")

        # nothing above reached a kernel: the context still trains and its
        # factors still match what they were before the refused calls
        P1, S1 = c.get_factors(0)
        assert np.array_equal(P, P1) and np.array_equal(S, S1)
        c.train(cs.make_opts(k=64, schedule=cs.GRID, num_workers=64), 0, 2)
        assert np.isfinite(c.test_rmse(1.0, 3.0))


def test_batch_refusals_leave_the_batch_usable(gpu):
    """Batched training refuses what it does not run (schedules other than
    Hogwild and serial, coherent mode, an unknown record order, k outside
    {32, 64, 128}) before any kernel, then still trains."""
    import cumf_sgd_amd as cs

    u, v, r, tu, tv, tr, m, n = _problem()
    with cs.Batch(cs.BF16) as b:
        _err(lambda: b.add(u, v, r, m, n, 48, 0.05), ERR_ARG, "k=48")
        b.add(u, v, r, m, n, 64, 0.05, tu=tu, tv=tv, tr=tr, seed=3)
        b.add(u[:5000], v[:5000], r[:5000], m, n, 32, 0.05, seed=3)
        _err(lambda: b.order(3), ERR_ARG, "bad order")
        _err(lambda: b.train(cs.make_opts(schedule=cs.STRATUM), 0, 1), ERR_ARG, "hogwild and serial")
        _err(lambda: b.train(cs.make_opts(schedule=cs.HOGWILD, coherent=1), 0, 1), ERR_ARG, "coherent")
        _err(lambda: b.train(cs.make_opts(schedule=cs.HOGWILD, batch_len=3), 0, 1), ERR_ARG, "even")
        P0, _ = b.get_factors(0, 0)
        b.order(cs.ORDER_COLUMNS)
        # the order moves records, never factors
        P1, _ = b.get_factors(0, 0)
        assert np.array_equal(P0, P1)
        b.train(cs.make_opts(schedule=cs.HOGWILD, seed=3), 0, 2)
        assert np.isfinite(b.test_rmse(0, 1.0, 3.0))
