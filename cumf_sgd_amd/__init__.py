"""Python binding of the cumf_sgd_amd C ABI (include/cumf_sgd_amd.h).

The library is the gfx950 build in this directory (libcumf_sgd_amd.so).  When
it is missing, the first import builds it (hipcc, build.py); there is no
fallback: if that build fails, importing the binding raises.
Arrays are numpy arrays (host) or objects exposing ``data_ptr()`` (torch
device tensors); both are passed as raw pointers.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

try:
    # PyTorch-ROCm ships its own HIP/HSA runtime (an unversioned
    # libamdhip64.so); the library links the system one.  Both can live in one
    # process only when the framework's runtime is loaded first, so import it
    # before the library when it is installed.
    import torch  # noqa: F401
except ImportError:  # the library itself does not need torch
    pass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcumf_sgd_amd.so")
CLI_PATH = os.path.join(_HERE, "cumf_sgd")
PREDICT_PATH = os.path.join(_HERE, "cumf_predict")

FP32, FP16, BF16, FP8_E4M3, BF16_SR = 0, 1, 2, 3, 4
PRECISIONS = {"fp32": FP32, "fp16": FP16, "bf16": BF16, "fp8": FP8_E4M3, "bf16sr": BF16_SR}
AUTO, HOGWILD, SERIAL, GRID, STRATUM, TILE = -1, 0, 1, 2, 3, 4
ORDER_ROWS, ORDER_COLUMNS = 1, 2
# mf-predict -e criteria (test/mf.h:30-40)
EVAL_RMSE, EVAL_MAE, EVAL_GKL, EVAL_LOGLOSS, EVAL_ACC = 0, 1, 2, 5, 6
EVAL_ROW_MPR, EVAL_COL_MPR, EVAL_ROW_AUC, EVAL_COL_AUC = 10, 11, 12, 13
SCHEDULES = {"hogwild": HOGWILD, "serial": SERIAL, "grid": GRID, "stratum": STRATUM, "tile": TILE}
ELEM_BYTES = {FP32: 4, FP16: 2, BF16: 2, FP8_E4M3: 1}
STORAGE_DTYPE = {FP32: np.float32, FP16: np.uint16, BF16: np.uint16, FP8_E4M3: np.uint8, BF16_SR: np.uint16}

NODE_DTYPE = np.dtype([("u", "<i4"), ("v", "<i4"), ("rate", "<f4")])


class CumfError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed (code {code}): {msg}")
        self.code = code


class TrainOpts(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32),
        ("num_workers", ctypes.c_int32),
        ("batch_len", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("lanes_per_sample", ctypes.c_int32),
        ("coherent", ctypes.c_int32),
        ("pipeline_depth", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("lambda_p", ctypes.c_float),
        ("lambda_q", ctypes.c_float),
        ("alpha", ctypes.c_float),
        ("beta", ctypes.c_float),
        ("lrate", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("stratum_rounds", ctypes.c_int32),
        ("tiles_in_flight", ctypes.c_int32),
    ]


class LaunchInfo(ctypes.Structure):
    _fields_ = [
        ("num_workers", ctypes.c_int32),
        ("lanes_per_sample", ctypes.c_int32),
        ("waves_per_block", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("batch_len", ctypes.c_int32),
        ("pipeline_depth", ctypes.c_int32),
        ("updates_per_epoch", ctypes.c_int64),
    ]


class Param(ctypes.Structure):
    """Mirror of Parameter (singleGPU/sgd.h:26-46)."""

    _fields_ = [(f, ctypes.c_int32) for f in
                ("k", "num_workers", "u_grid", "v_grid", "x_grid", "y_grid", "ux", "vy", "num_iters", "gpu")] + [
        (f, ctypes.c_float) for f in ("lambda_p", "lambda_q", "lrate", "alpha", "beta")]


class Node(ctypes.Structure):
    _fields_ = [("u", ctypes.c_int32), ("v", ctypes.c_int32), ("rate", ctypes.c_float)]


def _build_library() -> None:
    """Compile the library from its sources (build.py, loaded by path: going
    through this package would recurse into this import)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("_cumf_build", os.path.join(_HERE, "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    b.build_library()


def _load():
    if not os.path.exists(LIB_PATH):
        try:
            _build_library()
        except Exception as e:
            raise ImportError(f"{LIB_PATH} is not built and building it failed: {e}") from e
    lib = ctypes.CDLL(LIB_PATH)
    P, I32, I64, U64, F32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
    D = ctypes.POINTER(ctypes.c_double)
    sig = {
        "cumf_abi_version": ([], ctypes.c_int),
        "cumf_last_error": ([], ctypes.c_char_p),
        "cumf_create": ([ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
        "cumf_destroy": ([P], ctypes.c_int),
        "cumf_sync": ([P], ctypes.c_int),
        "cumf_set_ratings": ([P, P, I64, I32, I32], ctypes.c_int),
        "cumf_set_ratings_soa": ([P, P, P, P, I64, I32, I32], ctypes.c_int),
        "cumf_set_test": ([P, P, I64], ctypes.c_int),
        "cumf_set_test_soa": ([P, P, P, P, I64], ctypes.c_int),
        "cumf_get_ratings": ([P, P, P, P], ctypes.c_int),
        "cumf_alloc_ratings": ([P, I64, I32, I32, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P)],
                               ctypes.c_int),
        "cumf_commit_ratings": ([P], ctypes.c_int),
        "cumf_rating_moments": ([P, D, D], ctypes.c_int),
        "cumf_remap": ([P, P, P], ctypes.c_int),
        "cumf_scale_ratings": ([P, F32], ctypes.c_int),
        "cumf_partition": ([P, I32, I32, P], ctypes.c_int),
        "cumf_partition_strata": ([P, I32, P], ctypes.c_int),
        "cumf_partition_tiles": ([P, I32, I32, P], ctypes.c_int),
        "cumf_tile_rows": ([I32, I32], ctypes.c_int32),
        "cumf_get_tiles": ([P, ctypes.POINTER(I64), P, P, P], ctypes.c_int),
        "cumf_set_blocks": ([P, I32, I32, P], ctypes.c_int),
        "cumf_init_factors": ([P, I32, I32, U64], ctypes.c_int),
        "cumf_set_factors": ([P, I32, P, P], ctypes.c_int),
        "cumf_get_factors": ([P, I32, P, P], ctypes.c_int),
        "cumf_export_factors": ([P, I32, F32, P, P], ctypes.c_int),
        "cumf_train": ([P, ctypes.POINTER(TrainOpts), I32, I32, P, I64], ctypes.c_int),
        "cumf_test_rmse": ([P, F32, F32, D], ctypes.c_int),
        "cumf_test_eval": ([P, I32, F32, F32, D], ctypes.c_int),
        "cumf_read_ratings": ([P, ctypes.c_char_p, ctypes.POINTER(I64), ctypes.POINTER(I32), ctypes.POINTER(I32)],
                              ctypes.c_int),
        "cumf_read_test": ([P, ctypes.c_char_p, ctypes.POINTER(I64)], ctypes.c_int),
        "cumf_stream_ratings": ([P, P, I64, I32, I32, I32, P], ctypes.c_int),
        "cumf_set_model": ([P, I32, I32, I32, P, P], ctypes.c_int),
        "cumf_kernel_time": ([P, D, ctypes.POINTER(I64), ctypes.POINTER(I64)], ctypes.c_int),
        "cumf_reset_kernel_time": ([P], ctypes.c_int),
        "cumf_autotune": ([P, ctypes.POINTER(TrainOpts), ctypes.POINTER(TrainOpts), D], ctypes.c_int),
        "cumf_last_launch": ([P, ctypes.POINTER(LaunchInfo)], ctypes.c_int),
        "cumf_sgd_update_k128": ([ctypes.POINTER(Param), I32, I32, I64, I64, P, P, P, P, P, P, U64], ctypes.c_int),
        "cumf_sgd_train": ([ctypes.POINTER(Param), ctypes.POINTER(TrainOpts), P, I64, I32, I32, I32, U64, P, I64,
                            P, P, P, P], ctypes.c_int),
        "cumf_gen_random_map": ([ctypes.c_uint32, I32, P], ctypes.c_int),
        "cumf_collect_data": ([P, I64, ctypes.POINTER(F32), ctypes.POINTER(F32)], ctypes.c_int),
        "cumf_grid_problem": ([P, I64, I32, I32, I32, I32, P, P, ctypes.POINTER(I64), ctypes.POINTER(I64)],
                              ctypes.c_int),
        "cumf_read_problem": ([ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(I64), ctypes.POINTER(I32),
                               ctypes.POINTER(I32)], ctypes.c_int),
        "cumf_free_host": ([P], None),
        "cumf_save_model": ([ctypes.c_char_p, I32, I32, I32, F32, P, P], ctypes.c_int),
        "cumf_load_model": ([ctypes.c_char_p, ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I32),
                             ctypes.POINTER(F32), ctypes.POINTER(P), ctypes.POINTER(P)], ctypes.c_int),
        "cumf_batch_create": ([ctypes.c_int, I32, ctypes.POINTER(P)], ctypes.c_int),
        "cumf_batch_destroy": ([P], ctypes.c_int),
        "cumf_batch_add": ([P, P, P, P, I64, I32, I32, I32, F32, F32, P, P, P, I64, U64, ctypes.POINTER(I32)],
                           ctypes.c_int),
        "cumf_batch_order": ([P, I32], ctypes.c_int),
        "cumf_batch_train": ([P, ctypes.POINTER(TrainOpts), I32, I32], ctypes.c_int),
        "cumf_batch_test_rmse": ([P, I32, F32, F32, D], ctypes.c_int),
        "cumf_batch_test_eval": ([P, I32, I32, F32, F32, D], ctypes.c_int),
        "cumf_batch_autotune": ([P, ctypes.POINTER(TrainOpts), ctypes.POINTER(TrainOpts), D], ctypes.c_int),
        "cumf_batch_get_factors": ([P, I32, I32, P, P], ctypes.c_int),
        "cumf_batch_kernel_time": ([P, D, ctypes.POINTER(I64), ctypes.POINTER(I64)], ctypes.c_int),
        "cumf_order_blocks": ([P, I32], ctypes.c_int),
        "cumf_default_schedule": ([P, I32, I32, ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I32)],
                                  ctypes.c_int),
        "cumf_train_noalloc": ([ctypes.c_int, P, ctypes.POINTER(TrainOpts), I32, P, P, P, I64, I32, I32, P, P, P, P,
                                I32, I32], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def _ptr(a) -> int | None:
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    raise TypeError(f"unsupported array type {type(a)}")


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise CumfError(rc, where, lib.cumf_last_error().decode(errors="replace"))


def last_error() -> str:
    return lib.cumf_last_error().decode(errors="replace")


def lr_schedule(alpha: float, beta: float, lrate: float, n: int) -> np.ndarray:
    """lr[i] = float(alpha/(1+beta*i^1.5)+lrate) (singleGPU/sgd_kernel.cu:144-149)."""
    a, b, r = (float(np.float32(x)) for x in (alpha, beta, lrate))
    return np.array([a / (1 + b * i ** 1.5) + r for i in range(n)], dtype=np.float64).astype(np.float32)


@dataclass
class KernelTime:
    ms: float
    launches: int
    updates: int


class Context:
    """A device context: ratings, factors and training on one GPU."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib.cumf_create(device, ctypes.byref(h)), "cumf_create")
        self._h = h
        self.m = self.n = 0
        self.nnz = 0
        self.k = 0
        self.precision = None

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.cumf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- ratings
    def set_ratings(self, u, v, r, m: int, n: int) -> None:
        nnz = len(u)
        _check(lib.cumf_set_ratings_soa(self._h, _ptr(u), _ptr(v), _ptr(r), nnz, m, n), "cumf_set_ratings_soa")
        self.m, self.n, self.nnz = m, n, nnz

    def set_ratings_nodes(self, nodes: np.ndarray, m: int, n: int) -> None:
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        _check(lib.cumf_set_ratings(self._h, _ptr(nodes), len(nodes), m, n), "cumf_set_ratings")
        self.m, self.n, self.nnz = m, n, len(nodes)

    def alloc_ratings(self, nnz: int, m: int, n: int):
        """The context's own device rating arrays, as torch tensors that alias
        them (fill, then call commit_ratings)."""
        import torch

        pu, pv, pr = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib.cumf_alloc_ratings(self._h, nnz, m, n, ctypes.byref(pu), ctypes.byref(pv), ctypes.byref(pr)),
               "cumf_alloc_ratings")
        self.m, self.n, self.nnz = m, n, nnz

        class _Dev:
            def __init__(self, ptr, typestr):
                self.__cuda_array_interface__ = {"shape": (nnz,), "typestr": typestr, "data": (ptr, False),
                                                 "version": 3, "strides": None}

        return (torch.as_tensor(_Dev(pu.value, "<i4"), device="cuda"),
                torch.as_tensor(_Dev(pv.value, "<i4"), device="cuda"),
                torch.as_tensor(_Dev(pr.value, "<f4"), device="cuda"))

    def commit_ratings(self) -> None:
        _check(lib.cumf_commit_ratings(self._h), "cumf_commit_ratings")

    def set_test(self, u, v, r) -> None:
        _check(lib.cumf_set_test_soa(self._h, _ptr(u), _ptr(v), _ptr(r), len(u)), "cumf_set_test_soa")

    def get_ratings(self):
        u = np.empty(self.nnz, np.int32)
        v = np.empty(self.nnz, np.int32)
        r = np.empty(self.nnz, np.float32)
        _check(lib.cumf_get_ratings(self._h, _ptr(u), _ptr(v), _ptr(r)), "cumf_get_ratings")
        return u, v, r

    def rating_moments(self):
        s, s2 = ctypes.c_double(), ctypes.c_double()
        _check(lib.cumf_rating_moments(self._h, ctypes.byref(s), ctypes.byref(s2)), "cumf_rating_moments")
        return s.value, s2.value

    def remap(self, p_map: np.ndarray, q_map: np.ndarray) -> None:
        _check(lib.cumf_remap(self._h, _ptr(np.ascontiguousarray(p_map, np.int32)),
                              _ptr(np.ascontiguousarray(q_map, np.int32))), "cumf_remap")

    def scale_ratings(self, s: float) -> None:
        _check(lib.cumf_scale_ratings(self._h, float(s)), "cumf_scale_ratings")

    def partition(self, ux: int, vy: int) -> np.ndarray:
        off = np.zeros(ux * vy + 1, np.int64)
        _check(lib.cumf_partition(self._h, ux, vy, _ptr(off)), "cumf_partition")
        return off

    def partition_strata(self, s: int) -> np.ndarray:
        off = np.zeros(s * s + 1, np.int64)
        _check(lib.cumf_partition_strata(self._h, s, _ptr(off)), "cumf_partition_strata")
        return off

    def partition_tiles(self, s: int, tile_rows: int = 0) -> np.ndarray:
        """cumf_partition_tiles: s x s strata, each block's records grouped by
        row tile and ordered by column inside a tile; returns the block
        offsets."""
        off = np.zeros(s * s + 1, np.int64)
        _check(lib.cumf_partition_tiles(self._h, s, tile_rows, _ptr(off)), "cumf_partition_tiles")
        return off

    def get_tiles(self):
        """(rec_off[count + 1], row_begin[count], row_count[count]) of the tile partition."""
        n = ctypes.c_int64()
        _check(lib.cumf_get_tiles(self._h, ctypes.byref(n), None, None, None), "cumf_get_tiles")
        off = np.zeros(n.value + 1, np.int64)
        r0 = np.zeros(n.value, np.int32)
        rc = np.zeros(n.value, np.int32)
        _check(lib.cumf_get_tiles(self._h, ctypes.byref(n), _ptr(off), _ptr(r0), _ptr(rc)), "cumf_get_tiles")
        return off, r0, rc

    def order_blocks(self, by: int) -> None:
        """cumf_order_blocks: stable sort of each block's records by column
        (ORDER_COLUMNS) or row (ORDER_ROWS) id."""
        _check(lib.cumf_order_blocks(self._h, by), "cumf_order_blocks")

    def set_blocks(self, ux: int, vy: int, offsets: np.ndarray) -> None:
        _check(lib.cumf_set_blocks(self._h, ux, vy, _ptr(np.ascontiguousarray(offsets, np.int64))), "cumf_set_blocks")

    # ---- factors
    def init_factors(self, k: int, precision: int, seed: int) -> None:
        _check(lib.cumf_init_factors(self._h, k, precision, seed), "cumf_init_factors")
        self.k, self.precision = k, precision

    def get_factors(self, which: int):
        rows = self.m if which == 0 else self.n
        data = np.empty(rows * self.k, STORAGE_DTYPE[self.precision])
        scales = np.empty(rows, np.float32) if self.precision == FP8_E4M3 else None
        _check(lib.cumf_get_factors(self._h, which, _ptr(data), _ptr(scales)), "cumf_get_factors")
        return data.reshape(rows, self.k), scales

    def set_factors(self, which: int, data: np.ndarray, scales: np.ndarray | None = None) -> None:
        data = np.ascontiguousarray(data, STORAGE_DTYPE[self.precision])
        sc = None if scales is None else np.ascontiguousarray(scales, np.float32)
        _check(lib.cumf_set_factors(self._h, which, _ptr(data), _ptr(sc)), "cumf_set_factors")

    def export_factors(self, which: int, mult: float = 1.0, map_: np.ndarray | None = None) -> np.ndarray:
        rows = self.m if which == 0 else self.n
        out = np.empty((rows, self.k), np.float32)
        mp = None if map_ is None else np.ascontiguousarray(map_, np.int32)
        _check(lib.cumf_export_factors(self._h, which, float(mult), _ptr(mp), _ptr(out)), "cumf_export_factors")
        return out

    # ---- training
    def train(self, opts: TrainOpts, epoch_begin: int, epoch_end: int, block_order: np.ndarray | None = None) -> None:
        bo = None if block_order is None else np.ascontiguousarray(block_order, np.int32)
        _check(lib.cumf_train(self._h, ctypes.byref(opts), epoch_begin, epoch_end, _ptr(bo),
                              0 if bo is None else len(bo)), "cumf_train")

    def test_rmse(self, mult: float = 1.0, b: float = 0.0) -> float:
        out = ctypes.c_double()
        _check(lib.cumf_test_rmse(self._h, float(mult), float(b), ctypes.byref(out)), "cumf_test_rmse")
        return out.value

    def stream_ratings(self, nodes: np.ndarray, m: int, n: int, s: int) -> np.ndarray:
        """Out-of-core ratings for the stratum schedule: the 12-byte records
        (ids permuted, values scaled) stay in pinned host memory in s x s
        blocks and stream to the device one diagonal at a time during
        training.  Returns the block bounds."""
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        off = np.zeros(s * s + 1, np.int64)
        _check(lib.cumf_stream_ratings(self._h, _ptr(nodes), len(nodes), m, n, s, _ptr(off)), "cumf_stream_ratings")
        self.m, self.n, self.nnz = m, n, len(nodes)
        return off

    def read_ratings(self, path: str):
        """Stream the training file (or shards path0..path79) into the device
        arrays; returns (nnz, m, n) as read_problem computes them."""
        nnz, m, n = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib.cumf_read_ratings(self._h, path.encode(), ctypes.byref(nnz), ctypes.byref(m), ctypes.byref(n)),
               "cumf_read_ratings")
        self.m, self.n, self.nnz = m.value, n.value, nnz.value
        return nnz.value, m.value, n.value

    def read_test(self, path: str) -> int:
        nnz = ctypes.c_int64()
        _check(lib.cumf_read_test(self._h, path.encode(), ctypes.byref(nnz)), "cumf_read_test")
        return nnz.value

    def set_model(self, P: np.ndarray, Q: np.ndarray) -> None:
        """FP32 factors of any k for evaluation (cumf_set_model)."""
        P = np.ascontiguousarray(P, np.float32)
        Q = np.ascontiguousarray(Q, np.float32)
        _check(lib.cumf_set_model(self._h, P.shape[0], Q.shape[0], P.shape[1], _ptr(P), _ptr(Q)), "cumf_set_model")
        self.m, self.n, self.k, self.precision = P.shape[0], Q.shape[0], P.shape[1], FP32

    def test_eval(self, criterion: int, mult: float = 1.0, b: float = 0.0) -> float:
        """mf-predict's -e criterion (EVAL_RMSE, EVAL_MAE, EVAL_GKL, EVAL_LOGLOSS, EVAL_ACC)."""
        out = ctypes.c_double()
        _check(lib.cumf_test_eval(self._h, int(criterion), float(mult), float(b), ctypes.byref(out)), "cumf_test_eval")
        return out.value

    def kernel_time(self) -> KernelTime:
        ms = ctypes.c_double()
        la, up = ctypes.c_int64(), ctypes.c_int64()
        _check(lib.cumf_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(la), ctypes.byref(up)), "cumf_kernel_time")
        return KernelTime(ms.value, la.value, up.value)

    def prepare_default(self, k: int, precision: int, strata: int = 24) -> tuple[int, int]:
        """cumf_default_schedule, then the partition and record order it asks
        for (strata x strata blocks for STRATUM, the tile partition for
        TILE); returns (schedule, stratum_rounds) for the training options."""
        sched, rounds, order = self.default_schedule(k, precision)
        if sched == STRATUM:
            self.partition_strata(strata)
        elif sched == TILE:
            self.partition_tiles(8)
        elif order:
            self.partition(1, 1)
        if order:
            self.order_blocks(order)
        return sched, rounds

    def default_schedule(self, k: int, precision: int) -> tuple[int, int, int]:
        """cumf_default_schedule: (schedule, stratum_rounds, order) that a
        1x1 grid trains by default for these ratings (order 0: the records
        keep their order)."""
        out, rounds, order = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib.cumf_default_schedule(self._h, k, precision, ctypes.byref(out), ctypes.byref(rounds),
                                         ctypes.byref(order)), "cumf_default_schedule")
        return out.value, rounds.value, order.value

    def autotune(self, base: TrainOpts):
        """Fastest launch shape for this problem (factors are left unchanged);
        returns (opts, epoch milliseconds)."""
        best = TrainOpts()
        ms = ctypes.c_double()
        _check(lib.cumf_autotune(self._h, ctypes.byref(base), ctypes.byref(best), ctypes.byref(ms)), "cumf_autotune")
        return best, ms.value

    def reset_kernel_time(self) -> None:
        _check(lib.cumf_reset_kernel_time(self._h), "cumf_reset_kernel_time")

    def last_launch(self) -> LaunchInfo:
        info = LaunchInfo()
        _check(lib.cumf_last_launch(self._h, ctypes.byref(info)), "cumf_last_launch")
        return info

    def sync(self) -> None:
        _check(lib.cumf_sync(self._h), "cumf_sync")


class Batch:
    """Many independent problems trained together (cumf_batch_*)."""

    def __init__(self, precision: int, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib.cumf_batch_create(device, precision, ctypes.byref(h)), "cumf_batch_create")
        self._h = h
        self.precision = precision
        self.shapes = []

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.cumf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, u, v, r, m, n, k, lambda_p, lambda_q=None, tu=None, tv=None, tr=None, seed=1) -> int:
        idx = ctypes.c_int32()
        lq = lambda_p if lambda_q is None else lambda_q
        tn = 0 if tu is None else len(tu)
        _check(lib.cumf_batch_add(self._h, _ptr(u), _ptr(v), _ptr(r), len(u), m, n, k, float(lambda_p), float(lq),
                                  _ptr(tu), _ptr(tv), _ptr(tr), tn, seed, ctypes.byref(idx)), "cumf_batch_add")
        self.shapes.append((m, n, k))
        return idx.value

    def order(self, by: int) -> None:
        """Stable sort of every problem's training records by row
        (ORDER_ROWS) or column id (ORDER_COLUMNS)."""
        _check(lib.cumf_batch_order(self._h, by), "cumf_batch_order")

    def train(self, opts: "TrainOpts", epoch_begin: int, epoch_end: int) -> None:
        _check(lib.cumf_batch_train(self._h, ctypes.byref(opts), epoch_begin, epoch_end), "cumf_batch_train")

    def test_rmse(self, index: int, mult: float = 1.0, b: float = 0.0) -> float:
        out = ctypes.c_double()
        _check(lib.cumf_batch_test_rmse(self._h, index, float(mult), float(b), ctypes.byref(out)),
               "cumf_batch_test_rmse")
        return out.value

    def autotune(self, base: "TrainOpts"):
        """Fastest total waves, pipeline depth and (when base.batch_len is 0)
        window length for this batch (factors left unchanged); returns
        (opts, epoch milliseconds)."""
        best = TrainOpts()
        ms = ctypes.c_double()
        _check(lib.cumf_batch_autotune(self._h, ctypes.byref(base), ctypes.byref(best), ctypes.byref(ms)),
               "cumf_batch_autotune")
        return best, ms.value

    def test_eval(self, index: int, criterion: int, mult: float = 1.0, b: float = 0.0) -> float:
        out = ctypes.c_double()
        _check(lib.cumf_batch_test_eval(self._h, index, int(criterion), float(mult), float(b), ctypes.byref(out)),
               "cumf_batch_test_eval")
        return out.value

    def get_factors(self, index: int, which: int):
        m, n, k = self.shapes[index]
        rows = m if which == 0 else n
        data = np.empty(rows * k, STORAGE_DTYPE[self.precision])
        scales = np.empty(rows, np.float32) if self.precision == FP8_E4M3 else None
        _check(lib.cumf_batch_get_factors(self._h, index, which, _ptr(data), _ptr(scales)), "cumf_batch_get_factors")
        return data.reshape(rows, k), scales

    def kernel_time(self) -> KernelTime:
        ms = ctypes.c_double()
        la, up = ctypes.c_int64(), ctypes.c_int64()
        _check(lib.cumf_batch_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(la), ctypes.byref(up)),
               "cumf_batch_kernel_time")
        return KernelTime(ms.value, la.value, up.value)


NOALLOC_MAX_EPOCHS = 4028   # CUMF_NOALLOC_MAX_EPOCHS


def train_noalloc(opts: TrainOpts, precision: int, u, v, r, m: int, n: int, P, Q, epoch_begin: int, epoch_end: int,
                  ps=None, qs=None, device: int = 0, stream: int | None = None) -> None:
    """cumf_train_noalloc: epochs of the Hogwild or serialized schedule over
    ratings and factors in the caller's device memory (torch tensors or raw
    device addresses), updated in place; the library allocates nothing.
    `stream` is a hipStream_t address (None = the null stream)."""
    _check(lib.cumf_train_noalloc(device, stream, ctypes.byref(opts), precision, _ptr(u), _ptr(v), _ptr(r),
                                  int(u.numel() if hasattr(u, "numel") else len(u)), m, n, _ptr(P), _ptr(Q),
                                  _ptr(ps), _ptr(qs), epoch_begin, epoch_end), "cumf_train_noalloc")


def make_opts(k: int = 128, schedule: int = HOGWILD, num_workers: int = 0, lambda_p: float = 0.05,
              lambda_q: float | None = None, alpha: float = 0.08, beta: float = 0.3, lrate: float = 0.0,
              seed: int = 1, batch_len: int = 0, lanes_per_sample: int = 0, coherent: int = 0,
              pipeline_depth: int = 0, use_graph: int = 0, stratum_rounds: int = 0,
              tiles_in_flight: int = 0) -> TrainOpts:
    o = TrainOpts()
    o.k, o.schedule, o.num_workers = k, schedule, num_workers
    o.lambda_p = lambda_p
    o.lambda_q = lambda_p if lambda_q is None else lambda_q
    o.alpha, o.beta, o.lrate, o.seed = alpha, beta, lrate, seed
    o.batch_len, o.lanes_per_sample, o.coherent = batch_len, lanes_per_sample, coherent
    o.pipeline_depth = pipeline_depth
    o.use_graph = use_graph
    o.stratum_rounds = stratum_rounds
    o.tiles_in_flight = tiles_in_flight
    return o


def make_param(k: int = 128, num_workers: int = 12, u_grid: int = 1, v_grid: int = 1, x_grid: int = 1,
               y_grid: int = 1, num_iters: int = 20, gpu: int = 0, lambda_p: float = 0.05,
               lambda_q: float | None = None, lrate: float = 0.0, alpha: float = 0.08, beta: float = 0.3) -> Param:
    p = Param()
    if u_grid * v_grid == 1:
        x_grid = y_grid = 1
    p.k, p.num_workers = k, num_workers
    p.u_grid, p.v_grid, p.x_grid, p.y_grid = u_grid, v_grid, x_grid, y_grid
    p.ux, p.vy = u_grid * x_grid, v_grid * y_grid
    p.num_iters, p.gpu = num_iters, gpu
    p.lambda_p = lambda_p
    p.lambda_q = lambda_p if lambda_q is None else lambda_q
    p.lrate, p.alpha, p.beta = lrate, alpha, beta
    return p


# ---- host pipeline helpers ------------------------------------------------
def gen_random_map(seed: int, size: int) -> np.ndarray:
    out = np.empty(size, np.int32)
    _check(lib.cumf_gen_random_map(seed, size, _ptr(out)), "cumf_gen_random_map")
    return out


def collect_data(nodes: np.ndarray):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    a, s = ctypes.c_float(), ctypes.c_float()
    _check(lib.cumf_collect_data(_ptr(nodes), len(nodes), ctypes.byref(a), ctypes.byref(s)), "cumf_collect_data")
    return a.value, s.value


def grid_problem(nodes: np.ndarray, m: int, n: int, ux: int, vy: int):
    """In-place cycle-swap partition; returns (grid_size, offsets, u_seg, v_seg)."""
    if nodes.dtype != NODE_DTYPE or not nodes.flags["C_CONTIGUOUS"]:
        raise ValueError("nodes must be a contiguous NODE_DTYPE array")
    gs = np.zeros(ux * vy, np.int64)
    off = np.zeros(ux * vy + 1, np.int64)
    us, vs = ctypes.c_int64(), ctypes.c_int64()
    _check(lib.cumf_grid_problem(_ptr(nodes), len(nodes), m, n, ux, vy, _ptr(gs), _ptr(off), ctypes.byref(us),
                                 ctypes.byref(vs)), "cumf_grid_problem")
    return gs, off, us.value, vs.value


def read_problem(path: str):
    R = ctypes.c_void_p()
    nnz, m, n = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib.cumf_read_problem(path.encode(), ctypes.byref(R), ctypes.byref(nnz), ctypes.byref(m), ctypes.byref(n)),
           "cumf_read_problem")
    try:
        buf = (ctypes.c_char * (12 * nnz.value)).from_address(R.value) if nnz.value else b""
        nodes = np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()
    finally:
        lib.cumf_free_host(R)
    return nodes, m.value, n.value


def save_model(path: str, P: np.ndarray, Q: np.ndarray, b: float) -> None:
    P = np.ascontiguousarray(P, np.float32)
    Q = np.ascontiguousarray(Q, np.float32)
    _check(lib.cumf_save_model(path.encode(), P.shape[0], Q.shape[0], P.shape[1], float(b), _ptr(P), _ptr(Q)),
           "cumf_save_model")


def load_model(path: str):
    """The reference's model file (test/mf.cpp:3706-3777): returns (P, Q, b)."""
    m, n, k = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    b = ctypes.c_float()
    pp, qp = ctypes.c_void_p(), ctypes.c_void_p()
    _check(lib.cumf_load_model(path.encode(), ctypes.byref(m), ctypes.byref(n), ctypes.byref(k), ctypes.byref(b),
                               ctypes.byref(pp), ctypes.byref(qp)), "cumf_load_model")
    try:
        P = np.ctypeslib.as_array((ctypes.c_float * (m.value * k.value)).from_address(pp.value)).copy() \
            if m.value else np.zeros(0, np.float32)
        Q = np.ctypeslib.as_array((ctypes.c_float * (n.value * k.value)).from_address(qp.value)).copy() \
            if n.value else np.zeros(0, np.float32)
    finally:
        lib.cumf_free_host(pp)
        lib.cumf_free_host(qp)
    return P.reshape(m.value, k.value), Q.reshape(n.value, k.value), np.float32(b.value)


def sgd_train(param: Param, nodes: np.ndarray, m: int, n: int, precision: int = FP16, seed: int = 1,
              test: np.ndarray | None = None, opts: TrainOpts | None = None):
    """The full pipeline (cumf_sgd_train); returns (P, Q, b, rmse per epoch or None)."""
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    P = np.empty((m, param.k), np.float32)
    Q = np.empty((n, param.k), np.float32)
    b = ctypes.c_float()
    rmse = np.zeros(max(param.num_iters, 1), np.float64) if test is not None else None
    t = None if test is None else np.ascontiguousarray(test, NODE_DTYPE)
    _check(lib.cumf_sgd_train(ctypes.byref(param), None if opts is None else ctypes.byref(opts), _ptr(nodes),
                              len(nodes), m, n, precision, seed, _ptr(t), 0 if t is None else len(t), _ptr(P), _ptr(Q),
                              ctypes.byref(b), _ptr(rmse)), "cumf_sgd_train")
    return P, Q, b.value, rmse


def sgd_update_k128(param: Param, m: int, n: int, u_seg: int, v_seg: int, blocks: list[np.ndarray],
                    halfp: np.ndarray, halfq: np.ndarray, seed: int = 1):
    """Drop-in for sgd_update_k128: blocks are NODE_DTYPE arrays with local ids."""
    blocks = [np.ascontiguousarray(b, NODE_DTYPE) for b in blocks]
    ptrs = (ctypes.c_void_p * len(blocks))(*[_ptr(b) if len(b) else None for b in blocks])
    sizes = np.array([len(b) for b in blocks], np.int64)
    floatp = np.empty(halfp.shape, np.float32)
    floatq = np.empty(halfq.shape, np.float32)
    _check(lib.cumf_sgd_update_k128(ctypes.byref(param), m, n, u_seg, v_seg, ctypes.cast(ptrs, ctypes.c_void_p),
                                    _ptr(sizes), _ptr(halfp), _ptr(halfq), _ptr(floatp), _ptr(floatq), seed),
           "cumf_sgd_update_k128")
    return floatp, floatq
