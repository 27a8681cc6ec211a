"""CPU tests of the C ABI library and the host pipeline: every declared symbol
is exported, the reference's host algorithms are bit-exact, the file formats
match the reference's readers/writers, and failures are status codes."""
import os
import re
import subprocess

import numpy as np
import pytest

import cumf_sgd_amd as cs
import oracle
from cumf_sgd_amd import datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cumf_sgd_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cumf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", cs.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (cumf_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert cs.lib.cumf_abi_version() == 3


def test_library_has_gfx950_code():
    blob = open(cs.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_context_without_gpu_is_an_error_not_an_exit():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(cs.CumfError) as e:
        cs.Context(0)
    assert e.value.code == 2 and "device" in str(e.value)


def test_random_map_and_collect_data_bit_exact_with_oracle():
    for seed, size in [(0, 17), (2, 480_189), (99, 17_770)]:
        assert np.array_equal(cs.gen_random_map(seed, size), oracle.gen_random_map(seed, size))
    R = datagen.small(seed=4, m=2000, n=400, nnz=50_000)
    u, v, r, *_ = R.numpy()
    nodes = np.zeros(len(u), cs.NODE_DTYPE)
    nodes["u"], nodes["v"], nodes["rate"] = u, v, r
    assert cs.collect_data(nodes) == oracle.collect_data(nodes)


@pytest.mark.parametrize("ux,vy", [(1, 1), (2, 3), (8, 8), (13, 7)])
def test_grid_problem_bit_exact_with_oracle(ux, vy):
    rng = np.random.default_rng(ux * 10 + vy)
    m, n = 3001, 707
    nodes = np.zeros(40_000, cs.NODE_DTYPE)
    nodes["u"] = rng.integers(0, m, len(nodes))
    nodes["v"] = rng.integers(0, n, len(nodes))
    nodes["rate"] = rng.random(len(nodes))
    want, gs_w, off_w = oracle.grid_problem(nodes, m, n, ux, vy)
    got = nodes.copy()
    gs, off, us, vs = cs.grid_problem(got, m, n, ux, vy)
    assert np.array_equal(got, want) and np.array_equal(gs, gs_w) and np.array_equal(off, off_w)
    assert (us, vs) == oracle.grid_segments(m, n, ux, vy)


def test_grid_problem_rejects_out_of_range_ids():
    nodes = np.zeros(3, cs.NODE_DTYPE)
    nodes["u"] = [0, 5, 1]
    with pytest.raises(cs.CumfError) as e:
        cs.grid_problem(nodes, 4, 4, 2, 2)
    assert e.value.code == 1


def test_read_problem_single_file_and_truncation(tmp_path):
    rng = np.random.default_rng(1)
    u = rng.integers(0, 50, 1000).astype(np.int32)
    v = rng.integers(0, 20, 1000).astype(np.int32)
    r = rng.random(1000).astype(np.float32)
    p = tmp_path / "tr.bin"
    datagen.write_triplets(str(p), u, v, r)
    with open(p, "ab") as f:
        f.write(b"\x01\x02\x03\x04\x05")   # partial trailing record is dropped (sgd.cu:237)
    nodes, m, n = cs.read_problem(str(p))
    assert len(nodes) == 1000 and m == u.max() + 1 and n == v.max() + 1
    assert np.array_equal(nodes["u"], u) and np.array_equal(nodes["v"], v) and np.array_equal(nodes["rate"], r)


def test_read_problem_shards_with_first_offset_zero(tmp_path):
    rng = np.random.default_rng(2)
    parts = []
    for i in range(3):
        k = 100 + 37 * i
        u = rng.integers(0, 90, k).astype(np.int32)
        v = rng.integers(0, 30, k).astype(np.int32)
        r = rng.random(k).astype(np.float32)
        datagen.write_triplets(str(tmp_path / f"data{i}"), u, v, r)
        parts.append((u, v, r))
    nodes, m, n = cs.read_problem(str(tmp_path / "data"))
    assert len(nodes) == sum(len(p[0]) for p in parts)
    assert np.array_equal(nodes["u"], np.concatenate([p[0] for p in parts]))
    assert m == max(p[0].max() for p in parts) + 1 and n == max(p[1].max() for p in parts) + 1


def test_read_problem_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(cs.CumfError) as e:
        cs.read_problem(str(tmp_path / "nope.bin"))
    assert e.value.code == 5


def test_save_model_layout_matches_oracle_and_mf_predict(tmp_path):
    rng = np.random.default_rng(3)
    P = rng.normal(size=(40, 32)).astype(np.float32)
    Q = rng.normal(size=(25, 32)).astype(np.float32)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    cs.save_model(a, P, Q, 3.5)
    oracle.save_model(b, P, Q, np.float32(3.5))
    assert open(a, "rb").read() == open(b, "rb").read()
    if oracle.have_mf_predict():
        tu = rng.integers(0, 40, 500).astype(np.int32)
        tv = rng.integers(0, 25, 500).astype(np.int32)
        tr = rng.integers(1, 6, 500).astype(np.float32)
        t = str(tmp_path / "t.bin")
        datagen.write_triplets(t, tu, tv, tr)
        assert oracle.mf_predict_rmse(t, a) == f"{oracle.rmse(tu, tv, tr, P, Q, np.float32(3.5)):.4f}"


def test_save_model_unwritable_path_is_an_io_error(tmp_path):
    with pytest.raises(cs.CumfError) as e:
        cs.save_model(str(tmp_path / "no" / "such" / "dir" / "m"), np.zeros((1, 4), np.float32),
                      np.zeros((1, 4), np.float32), 0.0)
    assert e.value.code == 5


def _cli(*args, cwd=None):
    return subprocess.run([cs.CLI_PATH, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_usage_errors_return_one():
    assert _cli().returncode == 1
    out = _cli("-k")
    assert out.returncode == 1 and "need to specify" in out.stdout
    out = _cli("-s", "abc", "x")
    assert out.returncode == 1 and "-s should be followed by a number" in out.stdout


def test_cli_prints_reference_argument_block(tmp_path):
    """-s is rounded up to a multiple of 4 and x=y=1 when u*v == 1
    (singleGPU/main.cpp:139, 219-223); a missing training file is an I/O
    error, not exit(0)."""
    out = _cli("-s", "5", "-x", "3", "-k", "128", "-l", "0.05,0.02", str(tmp_path / "missing.bin"))
    assert "num_workers:8" in out.stdout
    assert "x_grid     :1" in out.stdout and "lambda_q   :0.02000" in out.stdout
    assert "model_path :missing.bin.model" in out.stdout
    assert out.returncode == 2 and "open failed" in out.stderr


def test_load_model_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(8)
    P = rng.normal(size=(23, 80)).astype(np.float32)   # any k: the reference's default is 80
    Q = rng.normal(size=(11, 80)).astype(np.float32)
    path = str(tmp_path / "m")
    cs.save_model(path, P, Q, 2.75)
    P2, Q2, b = cs.load_model(path)
    assert np.array_equal(P2, P) and np.array_equal(Q2, Q) and b == np.float32(2.75)
    with pytest.raises(cs.CumfError) as e:
        cs.load_model(str(tmp_path / "missing"))
    assert e.value.code == 5
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-4])
    with pytest.raises(cs.CumfError) as e:
        cs.load_model(path)
    assert e.value.code == 5 and "shorter" in str(e.value)


def _predict(*args):
    return subprocess.run([cs.PREDICT_PATH, *args], capture_output=True, text=True, timeout=60)


def test_predict_cli_usage_and_read_errors(tmp_path):
    out = _predict()
    assert out.returncode == 1 and "usage: cumf_predict" in out.stdout
    out = _predict("-e", "7", "a", "b")
    assert out.returncode == 1 and "unknown evaluation criterion" in out.stdout
    out = _predict("-e")
    assert out.returncode == 1 and "after -e" in out.stdout
    out = _predict("only_one")
    assert out.returncode == 1
    out = _predict(str(tmp_path / "t.bin"), str(tmp_path / "missing_model"))
    assert out.returncode == 2 and "cannot load model" in out.stderr
