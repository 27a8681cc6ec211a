"""ctypes wrapper of the CPU oracle (oracle.c) and the reference's CPU tools.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the reference to check against, never as
the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
MF_PREDICT = os.path.join(REF_DIR, "mf-predict")
MF_TRAIN = os.path.join(REF_DIR, "mf-train")

F32, F16, BF16, E4M3, BF16_SR = 0, 1, 2, 3, 4
STORAGE = {F32: np.float32, F16: np.uint16, BF16: np.uint16, E4M3: np.uint8, BF16_SR: np.uint16}

if not os.path.exists(LIB_PATH):
    # built on first use when missing; a failed build leaves the import failing
    _mk = subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    if _mk.returncode != 0 or not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built and `make -C oracle` failed:\n{_mk.stdout}")
_lib = ctypes.CDLL(LIB_PATH)

P = ctypes.c_void_p
I32, I64, U32, U64, F = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float


class SgdCfg(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("G", ctypes.c_int), ("prec", ctypes.c_int), ("batch_len", ctypes.c_int),
                ("workers", ctypes.c_int), ("problem", U32), ("seed", U64), ("lambda_p", F), ("lambda_q", F)]


_sig = {
    "orc_philox4x32_10": ([P, P, P], None),
    "orc_window_start": ([U64, U32, U32, U32, U32, U32, I64], I64),
    "orc_f16_to_f32": ([ctypes.c_uint16], F),
    "orc_f32_to_f16": ([F], ctypes.c_uint16),
    "orc_bf16_to_f32": ([ctypes.c_uint16], F),
    "orc_f32_to_bf16": ([F], ctypes.c_uint16),
    "orc_e4m3_to_f32": ([ctypes.c_uint8], F),
    "orc_f32_to_e4m3_rne": ([F], ctypes.c_uint8),
    "orc_f32_to_e4m3_sr": ([F, U32], ctypes.c_uint8),
    "orc_f32_to_bf16_sr": ([F, U32], ctypes.c_uint16),
    "orc_pow2_scale": ([F, F], F),
    "orc_init_factors": ([P, P, I64, ctypes.c_int, ctypes.c_int, U64, U32, ctypes.c_int], None),
    "orc_fp8_rescale": ([P, P, I64, ctypes.c_int], None),
    "orc_lr_table": ([F, F, F, ctypes.c_int, P], None),
    "orc_collect_data": ([P, I64, ctypes.POINTER(F), ctypes.POINTER(F)], None),
    "orc_gen_random_map": ([U32, I32, P], None),
    "orc_gen_inv_map": ([P, I32, P], None),
    "orc_shuffle_problem": ([P, I64, P, P], None),
    "orc_grid_segments": ([I32, I32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(I64), ctypes.POINTER(I64)], None),
    "orc_grid_problem": ([P, I64, I32, I32, ctypes.c_int, ctypes.c_int, P, P], None),
    "orc_scale_problem": ([P, I64, F, ctypes.c_int, ctypes.c_int, I64, I64], None),
    "orc_tree_dot": ([P, P, ctypes.c_int, ctypes.c_int], F),
    "orc_tree_dot_split": ([P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int], F),
    "orc_lane_elem": ([ctypes.c_int] * 5, ctypes.c_int),
    "orc_sgd_serial": ([ctypes.POINTER(SgdCfg), P, P, P, I64, P, P, P, P, P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "orc_sgd_block_windows": ([ctypes.POINTER(SgdCfg), U32, P, P, P, I64, P, P, P, P, P, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int], ctypes.c_int),
    "orc_sgd_serial_f64": ([ctypes.c_int, ctypes.c_int, U64, U32, ctypes.c_double, ctypes.c_double, P, P, P, I64, P, P,
                            P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "orc_widen": ([P, P, I64, ctypes.c_int, ctypes.c_int, P], None),
    "orc_scale_model": ([P, I64, ctypes.c_int, F], None),
    "orc_inv_shuffle": ([P, P, I32, ctypes.c_int], None),
    "orc_rmse": ([P, P, P, I64, P, P, I32, I32, ctypes.c_int, F], ctypes.c_double),
    "orc_eval": ([ctypes.c_int, P, P, P, I64, P, P, I32, I32, ctypes.c_int, F], ctypes.c_double),
    "orc_mpr_auc": ([P, P, P, I64, P, P, I32, I32, ctypes.c_int, F, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                     ctypes.POINTER(ctypes.c_double)], None),
    "orc_save_model": ([ctypes.c_char_p, I32, I32, I32, F, P, P], ctypes.c_int),
}
for _n, (_a, _r) in _sig.items():
    _f = getattr(_lib, _n)
    _f.argtypes = _a
    _f.restype = _r

NODE_DTYPE = np.dtype([("u", "<i4"), ("v", "<i4"), ("rate", "<f4")])


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    _lib.orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def window_start(seed, problem, tag, worker, epoch, batch, nnz):
    return _lib.orc_window_start(seed, problem, tag, worker, epoch, batch, nnz)


f16_to_f32 = _lib.orc_f16_to_f32
f32_to_f16 = _lib.orc_f32_to_f16
bf16_to_f32 = _lib.orc_bf16_to_f32
f32_to_bf16 = _lib.orc_f32_to_bf16
e4m3_to_f32 = _lib.orc_e4m3_to_f32
f32_to_e4m3_rne = _lib.orc_f32_to_e4m3_rne
f32_to_e4m3_sr = _lib.orc_f32_to_e4m3_sr
f32_to_bf16_sr = _lib.orc_f32_to_bf16_sr
pow2_scale = _lib.orc_pow2_scale


def init_factors(rows, k, prec, seed, which, problem=0):
    out = np.zeros(rows * k, STORAGE[prec])
    sc = np.zeros(rows, np.float32) if prec == E4M3 else None
    _lib.orc_init_factors(_p(out), _p(sc), rows, k, prec, seed, problem, which)
    return out.reshape(rows, k), sc


def fp8_rescale(data, scales):
    data = np.ascontiguousarray(data, np.uint8).copy()
    scales = np.ascontiguousarray(scales, np.float32).copy()
    _lib.orc_fp8_rescale(_p(data), _p(scales), data.shape[0], data.shape[1])
    return data, scales


def lr_table(alpha, beta, lrate, n):
    out = np.zeros(n, np.float32)
    _lib.orc_lr_table(alpha, beta, lrate, n, _p(out))
    return out


def collect_data(nodes):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    a, s = F(), F()
    _lib.orc_collect_data(_p(nodes), len(nodes), ctypes.byref(a), ctypes.byref(s))
    return a.value, s.value


def gen_random_map(seed, size):
    out = np.zeros(size, np.int32)
    _lib.orc_gen_random_map(seed, size, _p(out))
    return out


def gen_inv_map(mp):
    mp = np.ascontiguousarray(mp, np.int32)
    out = np.zeros_like(mp)
    _lib.orc_gen_inv_map(_p(mp), len(mp), _p(out))
    return out


def shuffle_problem(nodes, p_map, q_map):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE).copy()
    _lib.orc_shuffle_problem(_p(nodes), len(nodes), _p(np.ascontiguousarray(p_map, np.int32)),
                             _p(np.ascontiguousarray(q_map, np.int32)))
    return nodes


def grid_segments(m, n, ux, vy):
    a, b = I64(), I64()
    _lib.orc_grid_segments(m, n, ux, vy, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def grid_problem(nodes, m, n, ux, vy):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE).copy()
    gs = np.zeros(ux * vy, np.int64)
    off = np.zeros(ux * vy + 1, np.int64)
    _lib.orc_grid_problem(_p(nodes), len(nodes), m, n, ux, vy, _p(gs), _p(off))
    return nodes, gs, off


def scale_problem(nodes, scale, ux, vy, u_seg, v_seg):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE).copy()
    _lib.orc_scale_problem(_p(nodes), len(nodes), scale, ux, vy, u_seg, v_seg)
    return nodes


def tree_dot_split(p, q, G, half):
    """The dot product with lanes of two pieces of `half` elements, one in
    each half of the row (32-byte lanes)."""
    p = np.ascontiguousarray(p, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return _lib.orc_tree_dot_split(_p(p), _p(q), len(p), G, half)


def lane_elem(k, G, half, lane, i):
    return _lib.orc_lane_elem(k, G, half, lane, i)


def tree_dot(p, q, G):
    p = np.ascontiguousarray(p, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return _lib.orc_tree_dot(_p(p), _p(q), len(p), G)


def sgd_serial(u, v, r, P_, Q_, lr, epoch_begin, epoch_end, *, k, G, prec, lambda_p, lambda_q, seed, batch_len=128,
               ps=None, qs=None, problem=0):
    """Serialized single-worker epochs on copies of P_/Q_ (storage format)."""
    cfg = SgdCfg(k, G, prec, batch_len, 1, problem, seed, lambda_p, lambda_q)
    P2 = np.ascontiguousarray(P_).copy()
    Q2 = np.ascontiguousarray(Q_).copy()
    ps2 = None if ps is None else np.ascontiguousarray(ps, np.float32).copy()
    qs2 = None if qs is None else np.ascontiguousarray(qs, np.float32).copy()
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    r = np.ascontiguousarray(r, np.float32)
    lr = np.ascontiguousarray(lr, np.float32)
    _lib.orc_sgd_serial(ctypes.byref(cfg), _p(u), _p(v), _p(r), len(u), _p(P2), _p(Q2), _p(ps2), _p(qs2), _p(lr),
                        epoch_begin, epoch_end)
    return P2, Q2, ps2, qs2


def sgd_block_windows(u, v, r, P_, Q_, lr, epoch, tag, window_begin, windows, *, k, G, prec, lambda_p, lambda_q,
                      seed, batch_len=128, ps=None, qs=None, problem=0):
    """Windows [window_begin, window_begin + windows) of worker 0 of one rating
    block (tag = block id) in `epoch`, updating P_/Q_ (and FP8 scales) in
    place; they must be C-contiguous arrays of the storage format."""
    for a in (P_, Q_) + tuple(x for x in (ps, qs) if x is not None):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("factor arrays must be C-contiguous")
    cfg = SgdCfg(k, G, prec, batch_len, 1, problem, seed, lambda_p, lambda_q)
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    r = np.ascontiguousarray(r, np.float32)
    lr = np.ascontiguousarray(lr, np.float32)
    _lib.orc_sgd_block_windows(ctypes.byref(cfg), tag, _p(u), _p(v), _p(r), len(u), _p(P_), _p(Q_), _p(ps), _p(qs),
                               _p(lr), epoch, window_begin, windows)


def sgd_serial_f64(u, v, r, P_, Q_, lr, epoch_begin, epoch_end, *, k, lambda_p, lambda_q, seed, batch_len=128,
                   problem=0):
    P2 = np.ascontiguousarray(P_, np.float64).copy()
    Q2 = np.ascontiguousarray(Q_, np.float64).copy()
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    r = np.ascontiguousarray(r, np.float32)
    lr = np.ascontiguousarray(lr, np.float32)
    _lib.orc_sgd_serial_f64(k, batch_len, seed, problem, lambda_p, lambda_q, _p(u), _p(v), _p(r), len(u), _p(P2),
                            _p(Q2), _p(lr), epoch_begin, epoch_end)
    return P2, Q2


def widen(data, scales, prec):
    data = np.ascontiguousarray(data)
    rows, k = data.shape
    out = np.zeros((rows, k), np.float32)
    sc = None if scales is None else np.ascontiguousarray(scales, np.float32)
    _lib.orc_widen(_p(data), _p(sc), rows, k, prec, _p(out))
    return out


def scale_model(vec, scale):
    vec = np.ascontiguousarray(vec, np.float32).copy()
    _lib.orc_scale_model(_p(vec), vec.shape[0], vec.shape[1], scale)
    return vec


def inv_shuffle(vec, inv_map):
    vec = np.ascontiguousarray(vec, np.float32).copy()
    mp = np.ascontiguousarray(inv_map, np.int32).copy()
    _lib.orc_inv_shuffle(_p(vec), _p(mp), len(mp), vec.shape[1])
    return vec


def rmse(u, v, r, P_, Q_, b):
    P_ = np.ascontiguousarray(P_, np.float32)
    Q_ = np.ascontiguousarray(Q_, np.float32)
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    r = np.ascontiguousarray(r, np.float32)
    return _lib.orc_rmse(_p(u), _p(v), _p(r), len(u), _p(P_), _p(Q_), P_.shape[0], Q_.shape[0], P_.shape[1], b)


RMSE, MAE, GKL, LOGLOSS, ACC, ROW_MPR, COL_MPR, ROW_AUC, COL_AUC = 0, 1, 2, 5, 6, 10, 11, 12, 13


def mpr_auc(u, v, r, P_, Q_, b, transpose: bool):
    """calc_mpr_auc (test/mf.cpp:3914-4034): returns (MPR, AUC)."""
    P_ = np.ascontiguousarray(P_, np.float32)
    Q_ = np.ascontiguousarray(Q_, np.float32)
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    r = np.ascontiguousarray(r, np.float32)
    a, c = ctypes.c_double(), ctypes.c_double()
    _lib.orc_mpr_auc(_p(u), _p(v), _p(r), len(u), _p(P_), _p(Q_), P_.shape[0], Q_.shape[0], P_.shape[1], b,
                     int(transpose), ctypes.byref(a), ctypes.byref(c))
    return a.value, c.value


def evaluate(criterion, u, v, r, P_, Q_, b):
    """mf-predict's -e criterion on float factors (0 RMSE, 1 MAE, 2 GKL,
    5 LOGLOSS, 6 ACC, 10/11 row/column MPR, 12/13 row/column AUC)."""
    if criterion in (ROW_MPR, COL_MPR, ROW_AUC, COL_AUC):
        mpr, auc = mpr_auc(u, v, r, P_, Q_, b, criterion in (COL_MPR, COL_AUC))
        return mpr if criterion in (ROW_MPR, COL_MPR) else auc
    P_ = np.ascontiguousarray(P_, np.float32)
    Q_ = np.ascontiguousarray(Q_, np.float32)
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    r = np.ascontiguousarray(r, np.float32)
    return _lib.orc_eval(criterion, _p(u), _p(v), _p(r), len(u), _p(P_), _p(Q_), P_.shape[0], Q_.shape[0],
                         P_.shape[1], b)


def save_model(path, P_, Q_, b):
    P_ = np.ascontiguousarray(P_, np.float32)
    Q_ = np.ascontiguousarray(Q_, np.float32)
    return _lib.orc_save_model(path.encode(), P_.shape[0], Q_.shape[0], P_.shape[1], b, _p(P_), _p(Q_))


def have_mf_predict() -> bool:
    return os.access(MF_PREDICT, os.X_OK)


def mf_predict_rmse(test_path: str, model_path: str) -> str:
    """Run the reference's compiled evaluator: returns the 'RMSE = x.xxxx' value text."""
    return mf_predict_eval(test_path, model_path, RMSE)


_EVAL_LABEL = {RMSE: "RMSE = ", MAE: "MAE = ", GKL: "GKL = ", LOGLOSS: "LOGLOSS = ", ACC: "ACCURACY = ",
               ROW_MPR: "Row-wise MPR = ", COL_MPR: "Column-wise MPR = ", ROW_AUC: "Row-wise AUC = ",
               COL_AUC: "Colmn-wise AUC = "}


def mf_predict_eval(test_path: str, model_path: str, criterion: int) -> str:
    """Run the reference's compiled evaluator with -e criterion: returns the
    printed value text (4 decimals)."""
    out = subprocess.run([MF_PREDICT, "-e", str(criterion), test_path, model_path], capture_output=True, text=True,
                         check=True)
    for line in out.stdout.splitlines():
        if line.startswith(_EVAL_LABEL[criterion]):
            return line.split("=", 1)[1].strip()
    raise RuntimeError(f"no {_EVAL_LABEL[criterion]!r} line in mf-predict output:\n{out.stdout}")
