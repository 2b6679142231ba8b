#!/usr/bin/env python3
"""Regenerates tests/golden/golden.json from the CPU oracle.

For every case of fixtures.GOLDEN_CASES: digests of the inputs, the node
count, the digest of the canonical forest serialisation, the digest of the
predict output on the case's predict set and the training accuracy.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, os.path.join(ROOT, "oracle"), ROOT]

import fixtures  # noqa: E402
import pyoracle  # noqa: E402


def main():
    out = {}
    for name, case in fixtures.GOLDEN_CASES.items():
        X, y, Xp = fixtures.case_data(case)
        f = pyoracle.fit(X, y, **fixtures.fit_params(case))
        pred = f.predict(Xp)
        if case["task"] == 0:
            acc = float((f.predict(X).argmax(1) == y.astype(np.int64)).mean())
        else:
            acc = float(np.sqrt(np.mean((f.predict(X) / case["n_trees"] - y) ** 2)))
        out[name] = {
            "case": case,
            "X_sha256": fixtures.array_digest(X),
            "y_sha256": fixtures.array_digest(y),
            "n_nodes": int(f.n_nodes),
            "forest_sha256": fixtures.canonical_digest(fixtures.oracle_arrays(f), case["task"]),
            "predict_sha256": fixtures.array_digest(pred),
            "train_accuracy" if case["task"] == 0 else "train_rmse": round(acc, 6),
        }
        print(name, out[name]["n_nodes"], out[name]["forest_sha256"][:16], flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main()
