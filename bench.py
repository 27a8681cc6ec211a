"""Benchmark: rating-updates/s at k=128 and time-to-target test RMSE.

Workload (one GPU): the medium Netflix-shaped problem, 480,189 users x
17,770 items, 99,072,112 training + 1,408,395 test ratings (integers 1..5),
k = 128, lambda 0.05, alpha 0.08, beta 0.3, generated (seeded) on the host
and copied to HBM.  One step = one epoch of batch-Hogwild SGD over all
training ratings (the reference's update_per_sec = nnz * iters / time,
singleGPU/sgd_kernel.cu:244), by default under the library's schedule for
this problem and format (cumf_default_schedule; for the FP16 headline: the
blocks of 8 x 8 nnz-balanced strata split into tiles of 96 consecutive users,
each tile trained by one workgroup with its users' P rows in LDS, records
ordered by item inside a tile, one launch per diagonal with each diagonal's
blocks on distinct XCDs; --schedule stratum runs 24 x 24 strata with
item-ordered blocks, --schedule hogwild the reference's 1x1 mode).  Each
epoch's launches are captured once as a HIP graph and replayed (--no-graph
launches them one by one).  The launch shape is fixed (the library's
defaults for the format with 32-record windows, which is what cumf_autotune
picks on this workload), so that the profiled kernel is the timed one;
--autotune runs the autotuner first instead.
Inputs are resident in HBM before the timed region starts.

Printed: one JSON line with value (rating updates/s of the timed epochs),
roofline of the update kernel per level (L2 bytes, HBM bytes and VALU issue
over its HIP-event duration against each level's peak, from the committed
profile of the same kernel and launch shape; with graph replay the events
bracket each epoch's graph, so the average launch includes the short gaps
between launches), time-to-target RMSE for each factor format against the
FP32 GPU run (each format on its library default), and the reference's CPU trainer
(test/mf-train, built from the reference sources into oracle/_ref) timed on
the host's cores, or the CPU oracle when that build is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Aggregate L2 bandwidth of MI355X: 8 XCDs x 16 L2 channels x 128 bytes per
# clock x 2.4 GHz peak engine clock (16 channels of 128 B/clk per XCD L2: AMD
# CDNA 3 architecture white paper; CDNA 4 keeps the 4 MB, 16-channel L2 per
# XCD).  Every row load and store of an update is an L2 request, so this is
# the level that serves the algorithmic bytes.
L2_PEAK_GBS = 8 * 16 * 128 * 2.4
K = 128
# algorithmic bytes per update at k=128: 12-byte rating + P and Q rows read and
# written (+ the two FP32 row scales read, FP8)
BYTES_PER_UPDATE = {0: 12 + 4 * K * 4, 1: 12 + 4 * K * 2, 2: 12 + 4 * K * 2, 3: 12 + 4 * K * 1 + 2 * 4,
                    4: 12 + 4 * K * 2}
PREC_NAME = {0: "fp32", 1: "fp16", 2: "bf16", 3: "fp8_e4m3", 4: "bf16_sr"}
# How each format's write-back rounds, for reading its time to target: the
# round-to-nearest formats drop every step smaller than half a unit in the
# last place (FP16 from about 2^-12 of a value, BF16 from 2^-9), which moves
# their trajectory away from FP32's; the stochastically rounded ones keep
# every step in expectation.
WRITE_BACK = {0: "fp32 (exact)", 1: "round to nearest", 2: "round to nearest (drops steps below half a BF16 ulp: "
              "not FP32's trajectory)", 3: "stochastic rounding to E4M3 (3 mantissa bits)",
              4: "stochastic rounding (FP32's trajectory in expectation)"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def prepare(ctx, R, seed: int):
    """Device preprocessing of the reference pipeline (singleGPU/sgd.cu:644-679):
    statistics, id permutation, rating scaling.  Returns (scale, b, lambda)."""
    import cumf_sgd_amd as cs

    ctx.set_ratings(R.u, R.v, R.r, R.m, R.n)
    ctx.set_test(R.tu, R.tv, R.tr)
    s1, s2 = ctx.rating_moments()
    ex, ex2 = s1 / R.nnz, s2 / R.nnz
    ave = np.float32(ex)
    std = np.float32(np.sqrt(ex2 - ex * ex))
    scale = max(np.float32(1e-4), std)
    ctx.remap(cs.gen_random_map(seed, R.m), cs.gen_random_map(seed, R.n))
    ctx.scale_ratings(np.float32(1.0 / scale))
    return scale, np.float32(ave / std), np.float32(np.float32(0.05) / scale)


def run_epochs(ctx, prec, epochs, scale, b, lam, seed, workers, lanes, schedule, depth=0, batch_len=0,
               rmse_every=True, rounds=0, graph=0):
    """Fresh factors, `epochs` epochs; returns (per-epoch train seconds, rmse list)."""
    import cumf_sgd_amd as cs

    ctx.init_factors(K, prec, seed)
    opts = cs.make_opts(k=K, schedule=schedule, num_workers=workers, lambda_p=float(lam), seed=seed,
                        lanes_per_sample=lanes, pipeline_depth=depth, batch_len=batch_len, stratum_rounds=rounds,
                        use_graph=graph)
    fs = np.float32(np.sqrt(scale))
    times, rmses = [], []
    for ep in range(epochs):
        t0 = time.perf_counter()
        ctx.train(opts, ep, ep + 1)
        times.append(time.perf_counter() - t0)
        if rmse_every:
            rmses.append(ctx.test_rmse(fs, b))
    return times, rmses


def host_threads() -> tuple[int, dict]:
    """Threads for the CPU baseline: the CPUs this process may run on (its
    affinity mask), capped by OMP_NUM_THREADS when the host sets it (the
    share of the host's cores a one-GPU lease gets)."""
    aff = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(aff, env) if env > 0 else aff
    return max(1, threads), {"nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": env or None}


def cpu_baseline_port(R, sample: int, epochs: int):
    """No reference build on this host: the CPU oracle's serialized FP32 SGD
    (the GPU kernel's update rule, one thread) on a bounded sample."""
    import oracle

    u, v, r = (t[:sample].cpu().numpy() for t in (R.u, R.v, R.r))
    m, n = R.m, R.n
    P, _ = oracle.init_factors(m, K, oracle.F32, 2, 0)
    Q, _ = oracle.init_factors(n, K, oracle.F32, 2, 1)
    lr = oracle.lr_table(0.08, 0.3, 0.0, epochs + 4)
    t0 = time.perf_counter()
    oracle.sgd_serial(u, v, r, P, Q, lr, 0, epochs, k=K, G=16, prec=oracle.F32, lambda_p=0.05, lambda_q=0.05,
                      seed=2, batch_len=128)
    dt = time.perf_counter() - t0
    return {"value": sample * epochs / dt, "unit": "rating-updates/s", "cores": 1, "kind": "port",
            "sample": f"the CPU oracle's serialized FP32 SGD (this kernel's update rule, one thread) on the first "
                      f"{sample} training ratings, {epochs} epochs, k={K} ({dt:.1f} s); the reference's "
                      f"test/mf-train was not built on this host"}


def cpu_baseline(R, threads: int, sample: int, iters: int):
    """The reference's CPU trainer on a bounded sample of the training set."""
    import oracle

    if not os.access(oracle.MF_TRAIN, os.X_OK):
        return cpu_baseline_port(R, min(sample, 20_000_000), 1)
    from cumf_sgd_amd import datagen

    u, v, r = (t[:sample].cpu().numpy() for t in (R.u, R.v, R.r))
    with tempfile.TemporaryDirectory() as d:
        tr = os.path.join(d, "train.bin")
        datagen.write_triplets(tr, u, v, r)
        cmd = [oracle.MF_TRAIN, "--quiet", "-l2", "0.05", "-k", str(K), "-t", str(iters), "-s", str(threads),
               "-n", str(2 * threads + 1), "-r", "0.08", tr, os.path.join(d, "model")]
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    for line in out.stdout.splitlines():
        if line.startswith("perf:"):
            return {"value": float(line.split()[1]), "unit": "rating-updates/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"test/mf-train (LIBMF FPSGD, SSE+OpenMP) on "
                              f"{'all' if sample == R.nnz else 'the first'} {sample} training ratings "
                              f"of the same workload, {iters} iterations, k={K}, -s {threads}; its optimizer is "
                              f"AdaGrad-style FPSGD, not the GPU kernel's plain SGD"}
    return None


# VALU issue: 1024 SIMDs at 2.4 GHz, in SIMD-cycles per second.  A SIMD
# issues a wave64 VALU instruction per quad-cycle, and gfx950 issues two in
# one quad-cycle for some instruction pairs (SQ_ACTIVE_INST_VALU2 counts
# those quad-cycles; measured: v_mul_f32, v_add_u32, v_and_b32 and v_mov_b32
# at 1.6-1.8 wave64 instructions per CU per clock, v_fma_f32,
# v_fma_mix_f32 and v_cvt_pk_f16_f32 at 0.95-1.08), so instruction counts
# have no fixed peak; the level is the fraction of quad-cycles in which the
# SIMDs were issuing VALU work, a dual-issued pair counted once.
VALU_ISSUE_PEAK_G = 1024 * 2.4
# L2 requests: 8 XCDs x 16 channels, each accepting one request per clock
# (its 128 B/clock path serves one 128-byte read per clock), at 2.4 GHz, in
# requests per second.
L2_REQUEST_PEAK_G = 8 * 16 * 2.4


def tile_waves(workers: int, lanes: int) -> int:
    """Waves of a tile workgroup (the library's rule: 4, 8 or 16)."""
    w = -(-workers // (64 // lanes))
    return 4 if w <= 4 else (8 if w <= 8 else 16)


def profile_summary(prec: int, info, schedule: str, upl: float | None = None, order: str | None = None) -> dict | None:
    """The committed rocprofv3 counter summary (profiles/) of the update
    kernel on this workload and schedule, when it was measured on the same
    kernel and launch shape (window length and updates per launch, when the
    summary records them; the worker budget alone does not change a stratum
    launch once it exceeds the windows of a diagonal) and, when given, the
    same record order inside the blocks; None otherwise."""
    name = {0: "fp32", 1: "fp16", 2: "bf16", 3: "fp8", 4: "bf16sr"}[prec]
    tiled = schedule.startswith("tile")
    path = os.path.join(ROOT, "profiles", f"{'sgd_tile' if tiled else 'sgd_update'}_{name}_medium_{schedule}.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    # the medium workload's factors are under 4 GB: rows addressed with 32-bit offsets
    if tiled:
        shape = (f"sgd_tile_kernel<128, {info.lanes_per_sample}, {prec}, {info.pipeline_depth}, true, "
                 f"{tile_waves(info.num_workers, info.lanes_per_sample)}>")
    else:
        shape = f"sgd_update_kernel<128, {info.lanes_per_sample}, {prec}, false, {info.pipeline_depth}, true>"
    if shape not in d["kernel"]:
        return None
    if d.get("batch_len", 128) != info.batch_len:   # summaries without it used the default 128
        return None
    if upl is not None and "updates_per_launch" in d and abs(d["updates_per_launch"] - upl) > 1e-3 * upl:
        return None
    if order is not None and d.get("block_order", "input") != order:
        return None
    return d


def measured_traffic(prec: int, info, schedule: str, upl: float | None = None, order: str | None = None):
    """HBM-side bytes per launch of the update kernel (calibrated FETCH_SIZE +
    WRITE_SIZE) from the matching profile summary; None otherwise."""
    d = profile_summary(prec, info, schedule, upl, order)
    return None if d is None else d["traffic_bytes_per_launch"]


def roofline(prec: int, info, schedule: str, avg_launch_s: float, upl: float, order: str) -> dict:
    """Roofline of the update kernel, per level, over the launch's HIP-event
    duration:
      - l2: bytes the L2 serves.  In a stratum launch every row load and store
        of an update is an L2 request, so these are the algorithmic bytes (the
        12-byte record + P and Q rows read and written); a tile launch keeps
        its P rows in LDS, and its level is the measured request bytes of the
        matching profile (TCP_TCC_READ_REQ x 128 + TCP_TCC_WRITE_REQ x 64);
      - l2_requests: the profile's L2 requests (TCP_TCC_READ_REQ +
        TCP_TCC_WRITE_REQ, a 64-byte write counting as one) against one
        request per channel per clock;
      - hbm: the profile's calibrated FETCH_SIZE + WRITE_SIZE against 8 TB/s;
      - valu: the profile's VALU issue quad-cycles (SQ_ACTIVE_INST_VALU less
        SQ_ACTIVE_INST_VALU2, a dual-issued pair counted once) against the
        SIMDs' cycles.
    `bound` is the level with the largest fraction."""
    algo = BYTES_PER_UPDATE[prec] * upl / avg_launch_s / 1e9
    prof = profile_summary(prec, info, schedule, upl, order)
    tiled = schedule.startswith("tile")
    traffic = None if prof is None else prof["traffic_bytes_per_launch"]
    levels = {}
    if not tiled:
        levels["l2"] = {"achieved": algo, "peak": L2_PEAK_GBS, "frac": algo / L2_PEAK_GBS, "unit": "GB/s",
                        "bytes": "algorithmic (12-byte record + P and Q rows read and written)"}
    if prof is not None:
        req = prof["l2_request_bytes_per_launch"] / avg_launch_s / 1e9
        if tiled:
            levels["l2"] = {"achieved": req, "peak": L2_PEAK_GBS, "frac": req / L2_PEAK_GBS, "unit": "GB/s",
                            "bytes": "measured L2 requests (128-byte reads, 64-byte writes)"}
        else:
            levels["l2"]["measured_requests_gbs"] = req
        hbm = traffic / avg_launch_s / 1e9
        levels["hbm"] = {"achieved": hbm, "peak": HBM_PEAK_GBS, "frac": hbm / HBM_PEAK_GBS, "unit": "GB/s",
                         "bytes": "measured: calibrated FETCH_SIZE + WRITE_SIZE per launch"}
        nreq = prof["tcp_tcc_read_req_per_launch"] + prof["tcp_tcc_write_req_per_launch"]
        rq = nreq / avg_launch_s / 1e9
        levels["l2_requests"] = {"achieved": rq, "peak": L2_REQUEST_PEAK_G, "frac": rq / L2_REQUEST_PEAK_G,
                                 "unit": "G requests/s",
                                 "requests": f"measured TCP_TCC_READ_REQ + TCP_TCC_WRITE_REQ "
                                             f"({nreq / upl:.2f} per update)"}
        if "valu_dual_quad_cycles_per_launch" in prof:
            q = prof["valu_active_quad_cycles_per_launch"] - prof["valu_dual_quad_cycles_per_launch"]
            v = q * 4 / avg_launch_s / 1e9
            levels["valu"] = {"achieved": v, "peak": VALU_ISSUE_PEAK_G, "frac": v / VALU_ISSUE_PEAK_G,
                              "unit": "G SIMD-cycles/s",
                              "cycles": f"measured VALU issue quad-cycles x 4 (SQ_ACTIVE_INST_VALU - "
                                        f"SQ_ACTIVE_INST_VALU2; {prof['valu_insts_per_launch'] / upl:.1f} "
                                        f"wave-instructions per update)"}
    bound = max(levels, key=lambda k: levels[k]["frac"]) if levels else None
    kernel = "sgd_tile_kernel" if tiled else "sgd_update_kernel"
    return {"bound": bound, "achieved": levels[bound]["achieved"] if bound else None,
            "peak": levels[bound]["peak"] if bound else None, "unit": levels[bound]["unit"] if bound else None,
            "frac": levels[bound]["frac"] if bound else None, "traffic": traffic, "levels": levels,
            "effective_algorithmic_gbs": algo, "kernel": kernel, "avg_launch_ms": avg_launch_s * 1e3,
            "bytes_per_update": BYTES_PER_UPDATE[prec], "updates_per_launch": upl,
            "profile": None if prof is None else prof.get("source")}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed epochs first: the GPU reaches its steady epoch time after about 20")
    ap.add_argument("--precision", default="fp16", choices=["fp32", "fp16", "bf16", "bf16sr", "fp8"])
    ap.add_argument("--workers", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--schedule", default="default", choices=["default", "hogwild", "stratum", "tile"],
                    help="default: the library's choice for this problem and format (cumf_default_schedule: row "
                         "tiles for FP32, FP16 and BF16, strata for BF16-SR and FP8); stratum: s x s "
                         "conflict-free blocks, each on one XCD, balanced per diagonal; tile: the blocks of 8 x 8 "
                         "strata split into row tiles, each trained by one workgroup with its P rows in LDS; "
                         "hogwild: the reference's 1x1 mode")
    ap.add_argument("--strata", type=int, default=0,
                    help="s for the stratum and tile schedules (s x s blocks; 0 = 24 for strata, 8 for tiles)")
    ap.add_argument("--order", default="columns", choices=["columns", "rows", "input"],
                    help="record order inside each stratum block (cumf_order_blocks): by column (default, the "
                         "library's default for this workload), by row, or as generated")
    ap.add_argument("--rounds", type=int, default=0,
                    help="stratum schedule: passes over the diagonals per epoch (0 = library default)")
    ap.add_argument("--batch-len", type=int, default=32,
                    help="records per window (even, <= 128; 0 = the library's 128); 32 is what cumf_autotune "
                         "picks on this workload")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each epoch's kernels one by one instead of replaying them as a captured HIP graph")
    ap.add_argument("--rmse-epochs", type=int, default=20)
    ap.add_argument("--no-rmse", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="training ratings for the CPU baseline (-1 = all of them, 0 = skip)")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--depth", type=int, default=0, help="pipeline depth (2, 4, 8; 0 = library default)")
    ap.add_argument("--copy-probe-mb", type=int, default=0)
    ap.add_argument("--autotune", action="store_true",
                    help="pick lanes, depth, window length and workers with cumf_autotune (trial epochs before the "
                         "timed region) instead of the fixed shape: the library's defaults with --batch-len, which "
                         "is the autotuner's pick on this workload and keeps the profiled kernel shape fixed")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: multi-GPU sharding is designed, not built; "
                                   "run with --gpus 1"}))
        return 2

    import torch

    import cumf_sgd_amd as cs
    from cumf_sgd_amd import datagen

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        return 2
    prec = cs.PRECISIONS[args.precision]
    sched = cs.SCHEDULES.get(args.schedule)   # None: the library's default
    t0 = time.perf_counter()
    # generated on the host (deterministic across machines), then made resident in HBM
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    R = datagen.medium(seed=args.seed, device="cpu")
    log(f"generated {R.nnz} train + {R.tu.numel()} test ratings in {time.perf_counter() - t0:.1f}s")
    R = datagen.Ratings(R.m, R.n, *(t.to("cuda") for t in (R.u, R.v, R.r, R.tu, R.tv, R.tr)))
    torch.cuda.synchronize()
    if args.copy_probe_mb:
        # a plain copy of known size, for scaling the HBM byte counters of a profiled run
        src = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ctx = cs.Context(0)
    scale, b, lam = prepare(ctx, R, args.seed)
    default_rounds = 0
    if sched is None:
        ctx.init_factors(K, prec, args.seed)   # the tile rows follow the factor format
        sched, default_rounds = ctx.prepare_default(K, prec)
        order_of = ctx.default_schedule(K, prec)[2]
        args.order = {cs.ORDER_COLUMNS: "columns", cs.ORDER_ROWS: "rows"}.get(order_of, "input")
        args.strata = 8 if sched == cs.TILE else 24
    elif sched == cs.STRATUM:
        ctx.partition_strata(args.strata or 24)
        if args.order != "input":
            ctx.order_blocks(cs.ORDER_COLUMNS if args.order == "columns" else cs.ORDER_ROWS)
    elif sched == cs.TILE:
        ctx.init_factors(K, prec, args.seed)   # the tile rows follow the factor format
        ctx.partition_tiles(args.strata or 8)
    if not args.strata:
        args.strata = 8 if sched == cs.TILE else 24
    if not args.rounds:
        args.rounds = default_rounds
    sched_tag = {cs.STRATUM: f"stratum{args.strata}", cs.TILE: f"tile{args.strata}"}.get(sched, "hogwild")
    torch.cuda.synchronize()
    log(f"preprocessed: scale={scale:.5f} b={b:.5f}")

    # ---- throughput: warmup + K timed epochs (one step = one epoch)
    ctx.init_factors(K, prec, args.seed)
    opts = cs.make_opts(k=K, schedule=sched, num_workers=args.workers, lambda_p=float(lam), seed=args.seed,
                        lanes_per_sample=args.lanes, pipeline_depth=args.depth, stratum_rounds=args.rounds,
                        batch_len=args.batch_len, use_graph=int(not args.no_graph))
    if args.autotune:
        t1 = time.perf_counter()
        opts, tuned_ms = ctx.autotune(opts)
        log(f"autotuned in {time.perf_counter() - t1:.1f}s: lanes={opts.lanes_per_sample} depth={opts.pipeline_depth} "
            f"batch_len={opts.batch_len} workers={opts.num_workers} ({tuned_ms:.2f} ms/epoch)")
    # the epochs of each phase in one call, as the reference's training loop
    # runs all its iterations inside sgd_update_k128 (singleGPU/sgd_kernel.cu:212-248)
    if args.warmup:
        ctx.train(opts, 0, args.warmup)
    ctx.sync()
    torch.cuda.synchronize()
    ctx.reset_kernel_time()
    t0 = time.perf_counter()
    ctx.train(opts, args.warmup, args.warmup + args.steps)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = ctx.kernel_time()
    info = ctx.last_launch()
    value = R.nnz * args.steps / wall
    avg_launch_s = kt.ms / 1e3 / kt.launches
    upd_per_launch = kt.updates / kt.launches
    achieved = BYTES_PER_UPDATE[prec] * upd_per_launch / avg_launch_s / 1e9
    log(f"{args.steps} epochs in {wall:.3f}s -> {value / 1e9:.3f}e9 updates/s; kernel avg {avg_launch_s * 1e3:.3f} ms "
        f"per launch ({kt.launches} launches), {achieved:.0f} GB/s algorithmic; workers={info.num_workers} "
        f"lanes={info.lanes_per_sample}")

    # ---- time to target RMSE (evaluation excluded from the time): every
    # format trains the library's default schedule for this problem and that
    # format (cumf_default_schedule: FP32 row tiles in two passes, the 16-bit
    # and FP8 formats 24 x 24 strata with item-ordered blocks)
    ttt = None
    if not args.no_rmse:
        ttt = {}
        curves = {}
        shapes = {}
        for p in (0, 1, 2, 4, 3):
            ctx.init_factors(K, p, args.seed)
            sched_p, rounds_p = ctx.prepare_default(K, p)
            w, l, d, bl = (args.workers, args.lanes, args.depth, args.batch_len)
            if args.autotune:
                o_p, _ = ctx.autotune(cs.make_opts(k=K, schedule=sched_p, num_workers=w, lambda_p=float(lam),
                                                   seed=args.seed, lanes_per_sample=l, pipeline_depth=d,
                                                   stratum_rounds=rounds_p, batch_len=bl,
                                                   use_graph=int(not args.no_graph)))
                w, l, d, bl = o_p.num_workers, o_p.lanes_per_sample, o_p.pipeline_depth, o_p.batch_len
            times, rmses = run_epochs(ctx, p, args.rmse_epochs, scale, b, lam, args.seed, w, l, sched_p, d, bl,
                                      rounds=rounds_p, graph=int(not args.no_graph))
            info_p = ctx.last_launch()
            shapes[p] = {"schedule": {cs.STRATUM: "stratum24", cs.TILE: "tile8", cs.HOGWILD: "hogwild"}[sched_p],
                         "passes": rounds_p, "lanes_per_sample": info_p.lanes_per_sample,
                         "pipeline_depth": info_p.pipeline_depth, "batch_len": info_p.batch_len}
            curves[p] = (times, rmses)
            log(f"{PREC_NAME[p]}: rmse by epoch {[round(x, 4) for x in rmses]}")
        target = curves[0][1][-1] + 0.005
        for p, (times, rmses) in curves.items():
            hit = next((i for i, x in enumerate(rmses) if x <= target), None)
            ttt[PREC_NAME[p]] = {
                "final_rmse": round(rmses[-1], 5),
                "epochs_to_target": None if hit is None else hit + 1,
                "seconds_to_target": None if hit is None else round(float(sum(times[:hit + 1])), 4),
                "train_seconds_total": round(float(sum(times)), 4),
                "meets_target": hit is not None,
                "write_back": WRITE_BACK[p],
                "shape": shapes[p],
            }
        ttt["target_rmse"] = round(target, 5)
        ttt["target_rule"] = f"FP32 GPU test RMSE after {args.rmse_epochs} epochs + 0.005"
        gold = os.path.join(ROOT, "tests", "golden", "medium_oracle_rmse.json")
        if os.path.exists(gold) and args.seed == 2 and args.rmse_epochs <= 20:
            g = json.load(open(gold))
            ref64 = g["fp64"][args.rmse_epochs - 1]
            ttt["fp64_oracle_rmse"] = round(ref64, 5)
            # the same rule anchored on the serialized FP32 oracle at the same
            # epoch count instead of the FP32 GPU run
            otarget = g["fp32"][args.rmse_epochs - 1] + 0.005
            ttt["oracle_target_rmse"] = round(otarget, 5)
            ttt["oracle_target_rule"] = f"serialized FP32 CPU oracle after {args.rmse_epochs} epochs + 0.005"
            for p, (times, rmses) in curves.items():
                ttt[PREC_NAME[p]]["gap_to_fp64_oracle"] = round(rmses[-1] - ref64, 5)
                hit = next((i for i, x in enumerate(rmses) if x <= otarget), None)
                ttt[PREC_NAME[p]]["epochs_to_oracle_target"] = None if hit is None else hit + 1
                ttt[PREC_NAME[p]]["seconds_to_oracle_target"] = (None if hit is None
                                                                 else round(float(sum(times[:hit + 1])), 4))

    # ---- CPU baseline (rank 0, N=1 only)
    threads, host = host_threads()
    threads = args.cpu_threads or threads
    cpu = None
    if args.cpu_sample != 0:
        try:
            sample = R.nnz if args.cpu_sample < 0 else min(args.cpu_sample, R.nnz)
            cpu = cpu_baseline(R, threads, sample, args.cpu_iters)
            if cpu is not None:
                cpu.update(host)
        except Exception as e:  # the baseline never decides the GPU result
            log(f"cpu baseline failed: {e}")

    line = {
        "metric": "rating-updates/s at k=128 + time-to-target test RMSE, BF16/FP8 vs FP32 on GPU",
        "value": value,
        "unit": "rating-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the update's arithmetic; the factor storage format is in config
        "dtype": "f32",
        "data": "synthetic (seeded, generated on the host, resident in HBM)",
        "config": {"workload": "medium Netflix-shaped MF: 480189 x 17770, 99072112 train + 1408395 test, k=128, "
                               "lambda 0.05, alpha 0.08, beta 0.3, batch-Hogwild",
                   "precision": args.precision,
                   "factor_storage": {0: "f32", 1: "f16", 2: "bf16", 3: "fp8e4m3 + f32 row scale",
                                      4: "bf16 (stochastic rounding)"}[prec],
                   "schedule": sched_tag, "graph_replay": not args.no_graph,
                   "block_order": {cs.STRATUM: args.order, cs.TILE: "row tiles, columns within a tile"}.get(
                       sched, "input"),
                   "shape_source": "cumf_autotune" if args.autotune else
                   "fixed: library defaults with the given window length (cumf_autotune's pick on this workload)",
                   "workers": info.num_workers,
                   "lanes_per_sample": info.lanes_per_sample, "batch_len": info.batch_len,
                   "updates_per_epoch": info.updates_per_epoch},
        "roofline": roofline(prec, info, sched_tag, avg_launch_s, upd_per_launch,
                             {cs.STRATUM: args.order, cs.TILE: "row tiles, columns within a tile"}.get(sched, "input")),
        "time_to_target": ttt,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
