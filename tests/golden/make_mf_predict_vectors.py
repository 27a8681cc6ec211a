"""Known-answer vectors for the evaluation criteria, produced by the
reference's own evaluator (test/mf-predict, compiled from the reference's
sources into oracle/_ref/ by oracle/Makefile.ref).

For each case it writes a small model file and test file, runs
`mf-predict -e c` for c in 0 (RMSE), 1 (MAE), 2 (GKL), 5 (LOGLOSS), 6 (ACC)
(and 10-13, row/column MPR and AUC, on the case with distinct pairs),
and stores the inputs (npz) and the printed 4-decimal values
(mf_predict_vectors.json) beside this script.

    python tests/golden/make_mf_predict_vectors.py
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from cumf_sgd_amd import datagen  # noqa: E402

CRITERIA = {"rmse": 0, "mae": 1, "gkl": 2, "logloss": 5, "acc": 6}
RANKING = {"row_mpr": 10, "col_mpr": 11, "row_auc": 12, "col_auc": 13}


def cases():
    rng = np.random.default_rng(20261203)
    out = []
    # ratings 1..5, positive factors (GKL is defined), out-of-range ids
    m, n, k = 60, 40, 16
    P = np.abs(rng.normal(0.4, 0.2, (m, k))).astype(np.float32)
    Q = np.abs(rng.normal(0.4, 0.2, (n, k))).astype(np.float32)
    tu = rng.integers(-2, m + 3, 4000).astype(np.int32)
    tv = rng.integers(0, n + 2, 4000).astype(np.int32)
    tr = rng.integers(1, 6, 4000).astype(np.float32)
    out.append(("ratings_1to5", tu, tv, tr, P, Q, np.float32(3.6)))
    # +-1 labels, signed factors, one NaN row (predicts b)
    m, n, k = 45, 70, 32
    P = rng.normal(0, 0.3, (m, k)).astype(np.float32)
    Q = rng.normal(0, 0.3, (n, k)).astype(np.float32)
    P[3, 5] = np.nan
    tu = rng.integers(0, m, 3000).astype(np.int32)
    tv = rng.integers(0, n, 3000).astype(np.int32)
    tr = np.where(rng.random(3000) < 0.4, -1.0, 1.0).astype(np.float32)
    out.append(("labels_pm1", tu, tv, tr, P, Q, np.float32(0.25)))
    # wider value range, k = 8
    m, n, k = 30, 20, 8
    P = np.abs(rng.normal(1.0, 0.5, (m, k))).astype(np.float32)
    Q = np.abs(rng.normal(1.0, 0.5, (n, k))).astype(np.float32)
    tu = rng.integers(0, m, 2000).astype(np.int32)
    tv = rng.integers(0, n, 2000).astype(np.int32)
    tr = (rng.random(2000) * 20 + 0.5).astype(np.float32)
    out.append(("wide_k8", tu, tv, tr, P, Q, np.float32(7.0)))
    # distinct pairs with +-1 labels and columns past the model: the ranking
    # criteria (10-13) are defined on distinct positives
    m, n, k = 50, 35, 16
    P = rng.normal(0, 0.4, (m, k)).astype(np.float32)
    Q = rng.normal(0, 0.4, (n, k)).astype(np.float32)
    flat = rng.choice((m + 2) * (n + 3), 900, replace=False)
    tu = (flat // (n + 3)).astype(np.int32)
    tv = (flat % (n + 3)).astype(np.int32)
    tr = np.where(rng.random(900) < 0.6, 1.0, -1.0).astype(np.float32)
    out.append(("ranking", tu, tv, tr, P, Q, np.float32(0.05)))
    return out


def main():
    if not oracle.have_mf_predict():
        raise SystemExit("oracle/_ref/mf-predict is not built (make -C oracle ref)")
    res = {"tool": "test/mf-predict -e <c> <test> <model> (reference sources, oracle/Makefile.ref)", "cases": []}
    with tempfile.TemporaryDirectory() as d:
        for name, tu, tv, tr, P, Q, b in cases():
            model, test = os.path.join(d, "m"), os.path.join(d, "t.bin")
            oracle.save_model(model, P, Q, b)
            datagen.write_triplets(test, tu, tv, tr)
            entry = {"file": f"mf_predict_{name}.npz"}
            for key, c in CRITERIA.items():
                entry[key] = oracle.mf_predict_eval(test, model, c)
            if name == "ranking":
                for key, c in RANKING.items():
                    entry[key] = oracle.mf_predict_eval(test, model, c)
            np.savez_compressed(os.path.join(HERE, entry["file"]), tu=tu, tv=tv, tr=tr, P=P, Q=Q, b=b)
            res["cases"].append(entry)
            print(entry)
    json.dump(res, open(os.path.join(HERE, "mf_predict_vectors.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
