"""Host logic of bench.py that runs without a GPU: the algorithmic bytes per
update and the lookup of measured HBM traffic, which must only be reported
for the launch shape its profile was taken on."""
from types import SimpleNamespace

import bench


def _info(lanes=16, depth=2, batch_len=32, workers=516096):
    return SimpleNamespace(lanes_per_sample=lanes, pipeline_depth=depth, batch_len=batch_len, num_workers=workers)


def test_bytes_per_update_at_k128():
    # 12-byte rating record; P and Q rows of 128 elements read and written
    assert bench.BYTES_PER_UPDATE[0] == 12 + 2 * 2 * 128 * 4
    assert bench.BYTES_PER_UPDATE[1] == bench.BYTES_PER_UPDATE[2] == bench.BYTES_PER_UPDATE[4] == 12 + 2 * 2 * 128 * 2
    # FP8: one byte per element plus the two FP32 row scales read
    assert bench.BYTES_PER_UPDATE[3] == 12 + 2 * 2 * 128 + 2 * 4


def test_traffic_reported_only_for_the_profiled_shape():
    upl = 4128389.3333333335      # updates per launch of the profiled FP16 stratum-24 run
    t = bench.measured_traffic(1, _info(), "stratum24", upl)
    assert t is not None and 5e8 < t < 8e8
    assert bench.measured_traffic(1, _info(), "stratum24") == t
    assert bench.measured_traffic(1, _info(lanes=32), "stratum24", upl) is None
    assert bench.measured_traffic(1, _info(depth=4), "stratum24", upl) is None
    assert bench.measured_traffic(1, _info(batch_len=128), "stratum24", upl) is None
    assert bench.measured_traffic(1, _info(), "stratum24", upl * 1.01) is None
    assert bench.measured_traffic(1, _info(), "stratum7", upl) is None
    # the profile was taken with each block's records ordered by item
    assert bench.measured_traffic(1, _info(), "stratum24", upl, "columns") == t
    assert bench.measured_traffic(1, _info(), "stratum24", upl, "input") is None


def test_roofline_fractions_stay_below_one():
    """Every level of the bench's roofline at the profiled shape, from the
    committed summary's own timing: the algorithmic bytes against the
    aggregate L2 bandwidth, the measured HBM-side bytes against the HBM peak
    and the measured VALU instructions against the VALU issue peak are
    fractions of at most one, and the bound is the largest."""
    import json
    import os

    d = json.load(open(os.path.join(bench.ROOT, "profiles", "sgd_update_fp16_medium_stratum24.json")))
    r = bench.roofline(1, _info(), "stratum24", d["trace_avg_launch_ms_timed_region"] / 1e3, d["updates_per_launch"],
                       d["block_order"])
    assert r["traffic"] == d["traffic_bytes_per_launch"]
    assert set(r["levels"]) == {"l2", "l2_requests", "hbm", "valu"}
    assert all(0 < lv["frac"] <= 1 for lv in r["levels"].values())
    assert r["frac"] == max(lv["frac"] for lv in r["levels"].values())
    assert abs(bench.L2_PEAK_GBS - 39321.6) < 1e-6
    assert abs(bench.VALU_ISSUE_PEAK_G - 2457.6) < 1e-9
    assert abs(bench.L2_REQUEST_PEAK_G - 307.2) < 1e-9


def test_every_profile_gives_levels_at_most_one():
    """Every committed profile, over its own traced launch time: each level,
    the VALU one counted as issue quad-cycles (a dual-issued pair once), is a
    fraction of at most one, for every storage format and schedule."""
    import glob
    import json

    precs = {"fp32": 0, "fp16": 1, "bf16": 2, "fp8": 3, "bf16sr": 4}
    paths = sorted(glob.glob(bench.os.path.join(bench.ROOT, "profiles", "sgd_*_medium_*.json")))
    assert len(paths) >= 6
    for path in paths:
        d = json.load(open(path))
        base = bench.os.path.basename(path)[:-5].split("_")
        prec, sched = precs[base[2]], base[4]
        info = _info(lanes=d["lanes_per_sample"], workers=d["workers"])
        r = bench.roofline(prec, info, sched, d["trace_avg_launch_ms_timed_region"] / 1e3, d["updates_per_launch"],
                           d["block_order"])
        assert r["profile"] == d["source"], path
        assert {"l2", "l2_requests", "hbm", "valu"} <= set(r["levels"]), path
        assert all(0 < lv["frac"] <= 1 for lv in r["levels"].values()), (path, r["levels"])


def test_tile_roofline_uses_measured_l2_requests():
    """A tile launch's P rows stay in LDS: its L2 level is the profile's
    measured request bytes, not the algorithmic bytes; the profile is found
    only for the tile kernel's workgroup size."""
    import glob
    import json

    for path in glob.glob(bench.os.path.join(bench.ROOT, "profiles", "sgd_tile_*_medium_tile8.json")):
        d = json.load(open(path))
        prec = {"fp32": 0, "fp16": 1, "bf16": 2}[path.split("sgd_tile_")[1].split("_")[0]]
        info = _info(lanes=d["lanes_per_sample"], workers=d["workers"])
        r = bench.roofline(prec, info, "tile8", d["trace_avg_launch_ms_timed_region"] / 1e3, d["updates_per_launch"],
                           d["block_order"])
        assert r["kernel"] == "sgd_tile_kernel" and r["profile"] == d["source"]
        assert abs(r["levels"]["l2"]["achieved"] - d["l2_request_bytes_per_launch"] /
                   (d["trace_avg_launch_ms_timed_region"] / 1e3) / 1e9) < 1e-6
        assert all(0 < lv["frac"] <= 1 for lv in r["levels"].values())
        assert bench.profile_summary(prec, _info(lanes=d["lanes_per_sample"], workers=4 * d["workers"]), "tile8",
                                     d["updates_per_launch"], d["block_order"]) is None
