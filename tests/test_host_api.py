"""Runs the C++ API test program (tests/cpp/test_host_api.cc), written in
the shape of the reference's lib_ensembles tests, through pytest."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_host_api")
GOLDEN = os.path.join(HERE, "golden")


def _run(mode, out_dir=None):
    cmd = [BIN, mode, GOLDEN] + ([out_dir] if out_dir else [])
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_host_api_cpu():
    out = _run("cpu")
    assert "PASSED 0" in out


@pytest.mark.gpu
def test_host_api_gpu(tmp_path):
    out = _run("gpu", str(tmp_path))
    assert "PASSED 0" in out
    # The 400-tree random forest of the reference's spambase test, built and
    # predicted through the C++ API, equals the oracle's committed votes.
    import json
    gold = json.load(open(os.path.join(GOLDEN, "golden.json")))["spambase_rf"]
    votes = np.fromfile(tmp_path / "spambase_rf_votes.bin", dtype=np.int32)
    assert votes.size == 4601 * 2
    assert hashlib.sha256(votes.tobytes()).hexdigest() == gold["predict_sha256"]
