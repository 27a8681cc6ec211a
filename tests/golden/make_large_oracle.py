"""Serialized CPU oracle RMSE curves on the large sparse workload.

Runs the FP32 oracle (the kernel's arithmetic order, one worker) and the FP64
oracle for 6 epochs on exactly the data and preprocessing of
tests/test_gpu_large.py::test_large_sparse_formats (1,000,990 x 624,961,
host-generated with seed 3, srand(3) id maps, scale = std of the training
ratings, lambda 0.05, alpha 0.08, beta 0.3), and writes the test RMSE after
every epoch into large_oracle_rmse.json.  Takes 30-60 minutes on one core per
precision; run the two one after the other:
    python tests/golden/make_large_oracle.py [fp32|fp64]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cumf_sgd_amd import datagen  # noqa: E402

EPOCHS = 6
SEED = 3


def main(which: str) -> None:
    import torch

    torch.set_num_threads(4)
    R = datagen.large_sparse(seed=SEED, device="cpu")
    u, v, r, tu, tv, tr = R.numpy()
    m, n = R.m, R.n
    del R
    ex = r.astype(np.float64).sum() / len(r)
    ex2 = (r.astype(np.float64) ** 2).sum() / len(r)
    ave = np.float32(ex)
    std = np.float32(np.sqrt(ex2 - ex * ex))
    scale = max(np.float32(1e-4), std)
    pm = oracle.gen_random_map(SEED, m)
    qm = oracle.gen_random_map(SEED, n)
    u = pm[u]
    v = qm[v]
    tu = pm[tu]
    tv = qm[tv]
    rs = (r * np.float32(1.0 / scale)).astype(np.float32)
    del r
    lam = np.float32(np.float32(0.05) / scale)
    lr = oracle.lr_table(0.08, 0.3, 0.0, EPOCHS + 4)
    P, _ = oracle.init_factors(m, 128, 0, SEED, 0)
    Q, _ = oracle.init_factors(n, 128, 0, SEED, 1)
    fs = np.float32(np.sqrt(scale))
    b = np.float32(ave / std)
    curve = []
    if which == "fp64":
        P = P.astype(np.float64)
        Q = Q.astype(np.float64)
    for ep in range(EPOCHS):
        t0 = time.time()
        if which == "fp32":
            P, Q, _, _ = oracle.sgd_serial(u, v, rs, P, Q, lr, ep, ep + 1, k=128, G=64, prec=0, lambda_p=lam,
                                           lambda_q=lam, seed=SEED)
            Pe, Qe = (P * fs).astype(np.float32), (Q * fs).astype(np.float32)
        else:
            P, Q = oracle.sgd_serial_f64(u, v, rs, P, Q, lr, ep, ep + 1, k=128, lambda_p=float(lam),
                                         lambda_q=float(lam), seed=SEED)
            Pe, Qe = (P * float(fs)).astype(np.float32), (Q * float(fs)).astype(np.float32)
        curve.append(oracle.rmse(tu, tv, tr, Pe, Qe, b))
        print(f"{which} epoch {ep}: rmse {curve[-1]:.5f} ({time.time() - t0:.0f}s)", flush=True)
    path = os.path.join(HERE, "large_oracle_rmse.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    data[which] = curve
    data["workload"] = ("large sparse: 1000990 x 624961, 250272272 train + 2528003 test, seed 3 (host), k=128, "
                        "lambda 0.05, alpha 0.08, beta 0.3, serialized single worker, windows of 128")
    data["script"] = "tests/golden/make_large_oracle.py"
    json.dump(data, open(path, "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "fp32")
