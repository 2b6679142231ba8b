"""Level steps replayed from captured hipGraphs.

Once every open node of a level fits the device-sized tiers, the batched
build runs each further level as a replay of a captured graph whose kernels
read the level's counts from device memory (no host round trip between
levels; gpudte_algorithm.cc:145-253 synchronises the context at least three
times per 512-node chunk instead).  The forests must equal the build that
sizes every launch on the host from counts read back (BATCHED_HOST), the
plain per-node driver and the oracle, bit for bit; the profile must show the
replays.
"""
import numpy as np
import pytest

import pyoracle
from datamininglibs_amd import capi, synth
from parity import assert_same_forest

pytestmark = pytest.mark.gpu


def _fit(X, y, mode, **kw):
    capi.profile_reset()
    capi.profiling(True)
    try:
        f = capi.fit(X, y, capi.default_params(build_mode=mode, **kw))
    finally:
        capi.profiling(False)
    st = capi.profile_stats()
    return f, st.get("level_graph_replays", {"launches": 0})["launches"]


def _same(a, b):
    ca, cb = a.canonical(), b.canonical()
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k


CASES = [
    # algorithm, task, rows, features, classes, k, depth, min node, trees, tuning
    (capi.RANDOM_FOREST, capi.CLASSIFICATION, 30000, 16, 5, 4, 18, 5, 6, {}),  # chunk pipeline above, graphs below
    (capi.RANDOM_FOREST, capi.REGRESSION, 20000, 12, 0, 4, 16, 5, 4, {}),
    (capi.EXTRA_TREES, capi.CLASSIFICATION, 8000, 20, 3, 0, 1000, 5, 10, {}),  # deep trees: replays in batches
    (capi.EXTRA_TREES, capi.REGRESSION, 6000, 10, 0, 3, 25, 10, 6, {}),
    (capi.RANDOM_FOREST, capi.CLASSIFICATION, 1500, 8, 2, 3, 1000, 1, 8, {}),  # deep RF, every level in graphs
    # nodes above a low workgroup cut inside the graphs (one workgroup per
    # item through a scratch slab)
    (capi.RANDOM_FOREST, capi.CLASSIFICATION, 20000, 10, 4, 3, 20, 2, 4, {"cut_team": 64, "cut_block": 64}),
    (capi.RANDOM_FOREST, capi.REGRESSION, 12000, 8, 0, 3, 20, 2, 3, {"cut_team": 64, "cut_block": 64}),
    # roots above 32768 entries: host-sized levels partition them by chunks
    (capi.EXTRA_TREES, capi.CLASSIFICATION, 100000, 16, 4, 0, 14, 5, 3, {}),
    (capi.EXTRA_TREES, capi.REGRESSION, 60000, 12, 0, 4, 14, 5, 3, {}),
]


@pytest.mark.parametrize("case", CASES, ids=[f"alg{c[0]}-task{c[1]}-n{c[2]}-d{c[6]}-cut{c[9].get('cut_block', 0)}"
                                             for c in CASES])
def test_graph_levels_equal_host_levels_and_oracle(case):
    alg, task, n, f, c, k, depth, min_node, trees, tuning = case
    if task == capi.CLASSIFICATION:
        X, y = synth.gaussian_classes(seed=n + f, n=n, f=f, n_informative=min(f, 6), n_classes=c)
    else:
        X, y = synth.sine_regression(seed=n + f, n=n, f=f)
    kw = dict(algorithm=alg, task=task, n_trees=trees, n_features_per_node=k, max_depth=depth,
              min_node_size=min_node, n_classes=c, seed=77, **tuning)
    g, replays = _fit(X, y, capi.BUILD_BATCHED, **kw)
    h, replays_host = _fit(X, y, capi.BUILD_BATCHED_HOST, **kw)
    assert replays > 0 and replays_host == 0, (replays, replays_host)
    _same(g, h)
    o = pyoracle.fit(X, y, algorithm=alg, task=task, n_trees=trees, k=k, max_depth=depth, min_node_size=min_node,
                     n_classes=c, seed=77, n_threads=8)
    assert_same_forest(g, o)
    if alg == capi.EXTRA_TREES and task == capi.CLASSIFICATION:
        p = capi.fit(X, y, capi.default_params(build_mode=capi.BUILD_PLAIN, **kw))
        _same(g, p)


def test_graph_levels_reuse_scratch_across_fits():
    """Fits of other shapes between graph builds (the cached scratch is
    regrown and the graphs re-captured) still give the same forests."""
    X1, y1 = synth.gaussian_classes(seed=3, n=12000, f=10, n_informative=5, n_classes=4)
    X2, y2 = synth.spam_like(seed=4, n=3000, f=30)
    kw1 = dict(algorithm=capi.RANDOM_FOREST, n_trees=5, n_features_per_node=3, max_depth=20, n_classes=4, seed=5)
    kw2 = dict(algorithm=capi.EXTRA_TREES, n_trees=20, max_depth=1000, min_node_size=2, n_classes=2, seed=6)
    a1, _ = _fit(X1, y1, capi.BUILD_BATCHED, **kw1)
    a2, _ = _fit(X2, y2, capi.BUILD_BATCHED, **kw2)
    b1, _ = _fit(X1, y1, capi.BUILD_BATCHED, **kw1)
    b2, _ = _fit(X2, y2, capi.BUILD_BATCHED_HOST, **kw2)
    _same(a1, b1)
    _same(a2, b2)
