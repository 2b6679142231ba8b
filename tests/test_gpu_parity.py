"""HIP build + predict vs the CPU oracle, bit for bit, on seeded inputs."""
import numpy as np
import pytest

from datamininglibs_amd import capi, synth
import pyoracle
from parity import assert_same_forest, fit_both

pytestmark = pytest.mark.gpu


def small_classes(seed, n, f, c, informative=4):
    return synth.gaussian_classes(seed=seed, n=n, f=f, n_informative=informative, n_classes=c)


@pytest.mark.parametrize("algorithm", [capi.EXTRA_TREES, capi.RANDOM_FOREST])
def test_spam_like_small(algorithm):
    X, y = synth.spam_like(seed=0, n=600, f=20)
    g, o = fit_both(X, y, algorithm=algorithm, n_trees=5, k=4, max_depth=25, min_node_size=10, n_classes=2)
    assert_same_forest(g, o)
    Xp, _ = synth.spam_like(seed=1, n=300, f=20)
    assert np.array_equal(g.predict(Xp).astype(np.int64), o.predict(Xp).astype(np.int64))


@pytest.mark.parametrize("algorithm", [capi.EXTRA_TREES, capi.RANDOM_FOREST])
@pytest.mark.parametrize("c", [3, 10, 16, 24])
def test_multiclass(algorithm, c):
    # 16 classes fill the class columns of the kernels built for up to 16;
    # 24 classes take the kernels built for up to 32.
    X, y = small_classes(3, 3000, 16, c)
    g, o = fit_both(X, y, algorithm=algorithm, n_trees=3, k=5, max_depth=20, min_node_size=5, n_classes=c)
    assert_same_forest(g, o)


def test_rf_big_nodes_bucket_path():
    # 20k rows: root and first levels exceed the LDS chunk and take the
    # bucketed path.
    X, y = small_classes(5, 20000, 12, 4)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, n_trees=2, k=3, max_depth=12, min_node_size=5,
                    n_classes=4)
    assert_same_forest(g, o)


@pytest.mark.parametrize("c", [16, 24])
def test_rf_big_nodes_bucket_path_many_classes(c):
    # The bucketed path and the wave and workgroup tiers with every class
    # column of the kernels built for up to 16 classes, and with those of
    # the kernels built for up to 32.
    X, y = small_classes(9, 20000, 12, c, informative=6)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, n_trees=2, k=3, max_depth=12, min_node_size=5,
                    n_classes=c)
    assert_same_forest(g, o)


def test_rf_heavy_ties_histogram_path():
    # Few distinct values: buckets overflow the sort capacity and use the
    # LDS histogram.
    rng = np.random.default_rng(11)
    X = rng.integers(0, 6, size=(6, 12000)).astype(np.float32)
    y = ((X[0] + X[1] + rng.integers(0, 2, 12000)) % 3).astype(np.float32)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, n_trees=2, k=3, max_depth=8, min_node_size=5,
                    n_classes=3)
    assert_same_forest(g, o)


def test_rf_regression_big_nodes_ties_and_wide_keys():
    # 30k rows: the bucket evaluator of regression with narrow key spans
    # (features of few values: counting sort, overfull buckets) and wide
    # ones (continuous features: radix passes).
    rng = np.random.default_rng(21)
    n = 30000
    X = np.empty((6, n), dtype=np.float32)
    X[:3] = rng.integers(0, 40, size=(3, n)).astype(np.float32)
    X[3:] = rng.standard_normal((3, n)).astype(np.float32)
    y = (np.sin(X[0] / 7.0) + 0.3 * X[3] + 0.1 * rng.standard_normal(n)).astype(np.float32)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, task=capi.REGRESSION, n_trees=2, k=3, max_depth=10,
                    min_node_size=5)
    assert_same_forest(g, o)


@pytest.mark.parametrize("algorithm", [capi.EXTRA_TREES, capi.RANDOM_FOREST])
def test_regression(algorithm):
    X, y = synth.sine_regression(seed=2, n=5000, f=8)
    g, o = fit_both(X, y, algorithm=algorithm, task=capi.REGRESSION, n_trees=3, k=3, max_depth=12,
                    min_node_size=5)
    assert_same_forest(g, o)
    Xp, _ = synth.sine_regression(seed=3, n=500, f=8)
    assert np.array_equal(g.predict(Xp), o.predict(Xp))


def _params(algorithm, mode, **kw):
    return capi.default_params(algorithm, build_mode=mode, **kw)


@pytest.mark.parametrize("algorithm", [capi.EXTRA_TREES, capi.RANDOM_FOREST])
def test_plain_equals_batched_and_oracle_over_tree_batches(algorithm):
    """130 trees: the plain path builds them as batches of 100 + 30 and
    concatenates; the batched path builds all at once.  Both equal the
    oracle."""
    X, y = small_classes(7, 1500, 10, 3)
    kw = dict(n_trees=130, n_features_per_node=3, max_depth=14, min_node_size=4, n_classes=3, seed=77)
    gb = capi.fit(X, y, _params(algorithm, capi.BUILD_BATCHED, **kw))
    gp = capi.fit(X, y, _params(algorithm, capi.BUILD_PLAIN, **kw))
    cb, cp = gb.canonical(), gp.canonical()
    for k in cb:
        assert np.array_equal(cb[k], cp[k]), k
    # leaf counts kept from the build (batched) and from the merge of the
    # plain path's batches equal the count over the node records
    assert gb.info.n_leaves == gp.info.n_leaves == int((cb["feature"] < 0).sum())
    nb, pb = gb.export()
    npl, ppl = gp.export()
    assert np.array_equal(nb, npl) and np.array_equal(pb, ppl)
    o = pyoracle.fit(X, y, algorithm=algorithm, n_trees=130, k=3, max_depth=14, min_node_size=4, n_classes=3,
                     seed=77)
    assert_same_forest(gp, o)
    Xp, _ = small_classes(8, 400, 10, 3)
    assert np.array_equal(gp.predict(Xp).astype(np.int64), o.predict(Xp).astype(np.int64))


@pytest.mark.parametrize("f", [12, 250, 400])
def test_predict_row_staging_widths(f):
    """Classification predict stages a block's rows in LDS: 64 rows while
    features + classes <= 191, 32 rows while <= 383, else rows are read from
    HBM.  All three kernels, feature-major (predict) and row-major (serving),
    give the oracle's votes; 1000 rows leave a ragged last block for both
    row counts."""
    X, y = small_classes(31, 800, f, 7)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, n_trees=12, k=6, max_depth=10, min_node_size=3,
                    n_classes=7)
    Xp, _ = small_classes(32, 1000, f, 7)
    want = o.predict(Xp).astype(np.int64)
    assert np.array_equal(g.predict(Xp).astype(np.int64), want)
    srv = capi.Server(max_batch_rows=1024)
    try:
        slot = srv.load(g)
        got = srv.wait(srv.submit(slot, np.ascontiguousarray(Xp.T)))
        assert np.array_equal(np.asarray(got).astype(np.int64), want)
    finally:
        srv.close()


@pytest.mark.parametrize("f", [12, 250, 400])
def test_regression_predict_row_staging_widths(f):
    """Regression predict stages 64 rows while features + 32 <= 191, 32 rows
    while <= 383, else reads rows from HBM; every row's sum is formed in tree
    order in all three, so predict and serving equal the oracle bit for bit.
    40 trees span two chunks of 32."""
    X, y = synth.sine_regression(seed=33, n=800, f=f)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, task=capi.REGRESSION, n_trees=40, k=6, max_depth=8,
                    min_node_size=3)
    Xp, _ = synth.sine_regression(seed=34, n=1000, f=f)
    want = o.predict(Xp)
    assert np.array_equal(g.predict(Xp), want)
    srv = capi.Server(max_batch_rows=1024)
    try:
        slot = srv.load(g)
        got = srv.wait(srv.submit(slot, np.ascontiguousarray(Xp.T)))
        assert np.array_equal(np.asarray(got), want)
    finally:
        srv.close()


@pytest.mark.parametrize("algorithm", [capi.EXTRA_TREES, capi.RANDOM_FOREST])
def test_plain_regression(algorithm):
    X, y = synth.sine_regression(seed=4, n=3000, f=6)
    g, o = fit_both(X, y, algorithm=algorithm, task=capi.REGRESSION, n_trees=3, k=2, max_depth=10,
                    min_node_size=5, build_mode=capi.BUILD_PLAIN)
    assert_same_forest(g, o)


def test_cached_scratch_reuse_and_release():
    """Fits reuse the scratch the previous fit left cached (a larger fit,
    then a smaller one of another shape, then a release of the cache and the
    first fit again): every forest equals its oracle forest."""
    Xa, ya = small_classes(21, 20000, 12, 4)
    Xb, yb = small_classes(22, 3000, 7, 3)
    kwa = dict(algorithm=capi.RANDOM_FOREST, n_trees=2, k=3, max_depth=12, min_node_size=5, n_classes=4)
    kwb = dict(algorithm=capi.RANDOM_FOREST, n_trees=3, k=2, max_depth=10, min_node_size=5, n_classes=3)
    g, o = fit_both(Xa, ya, **kwa)
    assert_same_forest(g, o)
    g, o = fit_both(Xb, yb, **kwb)
    assert_same_forest(g, o)
    assert capi.release_cached_memory() > 0
    assert capi.release_cached_memory() == 0
    g, o = fit_both(Xa, ya, **kwa)
    assert_same_forest(g, o)


@pytest.mark.parametrize("tuning", [dict(cut_team=128), dict(cut_team=64, cut_block=512), dict(cut_block=1024),
                                    dict(cut_team=384), dict(bucket_target=448),
                                    dict(bucket_target=64), dict(bucket_target=512),
                                    dict(cut_team=256, cut_block=1536, bucket_target=192)])
def test_rf_tuning_does_not_change_the_forest(tuning):
    """Every tier split and bucket size gives the oracle's forest (20k rows:
    nodes in every tier, the chunk pipeline at the top levels)."""
    X, y = small_classes(5, 20000, 12, 4)
    kw = dict(n_trees=2, n_features_per_node=3, max_depth=12, min_node_size=5, n_classes=4, seed=7)
    g = capi.fit(X, y, capi.default_params(capi.RANDOM_FOREST, **kw, **tuning))
    o = pyoracle.fit(X, y, algorithm=capi.RANDOM_FOREST, n_trees=2, k=3, max_depth=12, min_node_size=5,
                     n_classes=4, seed=7)
    assert_same_forest(g, o)


@pytest.mark.parametrize("cut_ert", [64, 1024, 65536])
def test_ert_tuning_does_not_change_the_forest(cut_ert):
    X, y = small_classes(6, 12000, 10, 3)
    kw = dict(n_trees=3, n_features_per_node=3, max_depth=14, min_node_size=5, n_classes=3, seed=5)
    g = capi.fit(X, y, capi.default_params(capi.EXTRA_TREES, cut_ert=cut_ert, **kw))
    o = pyoracle.fit(X, y, algorithm=capi.EXTRA_TREES, n_trees=3, k=3, max_depth=14, min_node_size=5,
                     n_classes=3, seed=5)
    assert_same_forest(g, o)


@pytest.mark.parametrize("algorithm", [capi.EXTRA_TREES, capi.RANDOM_FOREST])
def test_autotune_picks_a_valid_setting(algorithm):
    X, y = small_classes(8, 30000, 16, 5)
    p = capi.default_params(algorithm, n_trees=40, n_features_per_node=4, max_depth=16, min_node_size=5,
                            n_classes=5, seed=3)
    Xd = capi.DeviceArray.from_numpy(X)
    tuned, secs = capi.autotune(Xd.ptr, X.shape[1], X.shape[0], y, p, trial_trees=4)
    assert secs > 0
    if algorithm == capi.RANDOM_FOREST:
        assert tuned.cut_team in (256, 384, 512)
        assert tuned.cut_block in (1536, 2048) and tuned.cut_block >= tuned.cut_team
        assert tuned.bucket_target in (256, 320, 384, 448)
    else:
        assert tuned.cut_ert in (1024, 4096, 16384, 65536)
    assert tuned.n_trees == 40 and tuned.seed == 3
    a = capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, tuned).canonical()
    b = capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, p).canonical()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    Xd.free()


@pytest.mark.parametrize("algorithm,task", [(capi.RANDOM_FOREST, capi.CLASSIFICATION),
                                            (capi.EXTRA_TREES, capi.CLASSIFICATION),
                                            (capi.RANDOM_FOREST, capi.REGRESSION)])
def test_forest_file_round_trip(tmp_path, algorithm, task):
    """Save, reload: the same canonical arrays, export and predictions; a
    flipped byte is rejected by the checksum."""
    if task == capi.REGRESSION:
        X, y = synth.sine_regression(seed=4, n=4000, f=6)
        kw = dict(task=task, n_classes=0)
    else:
        X, y = small_classes(9, 4000, 10, 4)
        kw = dict(n_classes=4)
    f = capi.fit(X, y, capi.default_params(algorithm, n_trees=7, max_depth=12, seed=11, **kw))
    path = tmp_path / "forest.dtef"
    f.save(path)
    g = capi.load_forest(path)
    a, b = f.canonical(), g.canonical()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    (na, pa), (nb, pb) = f.export(), g.export()
    assert np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert f.info.n_leaves == g.info.n_leaves == int((a["feature"] < 0).sum())
    Xp = X[:, :500]
    assert np.array_equal(f.predict(Xp), g.predict(Xp))
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    bad = tmp_path / "bad.dtef"
    bad.write_bytes(bytes(raw))
    with pytest.raises(capi.DteError):
        capi.load_forest(bad)
    f.close()
    g.close()


@pytest.mark.parametrize("task,nf", [(capi.CLASSIFICATION, 6), (capi.REGRESSION, 5), (capi.CLASSIFICATION, 250),
                                     (capi.REGRESSION, 250), (capi.CLASSIFICATION, 400)])
def test_out_of_bag_pass_matches_oracle(task, nf):
    """Out-of-bag pairs, plain and with a feature permuted per visit, equal the
    oracle's (counts exactly, squared errors to 1e-9); importance is the drop.
    Feature counts cover the 64-row and 32-row LDS-staged kernels and the one
    reading rows from HBM (above 383 features)."""
    if task == capi.REGRESSION:
        X, y = synth.sine_regression(seed=6, n=1500, f=nf)
        kw = dict(task=task, n_classes=0)
    else:
        X, y = small_classes(10, 1500, nf, 3)
        kw = dict(n_classes=3)
    p = capi.default_params(capi.RANDOM_FOREST, n_trees=5, n_features_per_node=2, max_depth=10, seed=21, **kw)
    Xd = capi.DeviceArray.from_numpy(X)
    g = capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, p)
    o = pyoracle.fit(X, y, algorithm=capi.RANDOM_FOREST, task=task, n_trees=5, k=2, max_depth=10,
                     n_classes=3 if task == capi.CLASSIFICATION else 0, seed=21)
    assert_same_forest(g, o)
    for perm in (-1, 0, 2):
        gp, gv = capi.oob_evaluate(g, Xd.ptr, y, X.shape[1], perm)
        op, ov = pyoracle.oob_evaluate(o, X, y, seed=21, permuted_feature=perm)
        assert gp == op, perm
        if task == capi.CLASSIFICATION:
            assert gv == ov, perm
        else:
            assert abs(gv - ov) <= 1e-9 * max(1.0, abs(ov)), (perm, gv, ov)
    base, imp = capi.permutation_importance(g, Xd.ptr, y, X.shape[1])
    assert imp.shape == (X.shape[0],) and np.isfinite(imp).all()
    g.close()
    Xd.free()


def test_out_of_bag_pairs_over_row_ranges():
    """2.2 M rows: the in-bag bits are built in three ranges of 2^20 rows.
    The out-of-bag pairs of tree t are the rows its bootstrap never drew, N
    minus the distinct rows at its root, so the pass must count exactly
    sum_t (N - root entries)."""
    X, y = small_classes(14, 2_200_000, 4, 3)
    n = X.shape[1]
    p = capi.default_params(capi.RANDOM_FOREST, n_trees=3, n_features_per_node=2, max_depth=3, n_classes=3, seed=5)
    Xd = capi.DeviceArray.from_numpy(X)
    g = capi.fit_device(Xd.ptr, n, X.shape[0], y, p)
    c = g.canonical()
    roots = c["tree_start"][:-1]
    want = int(sum(n - int(c["n_entries"][r]) for r in roots))
    pairs, correct = capi.oob_evaluate(g, Xd.ptr, y, n, -1)
    assert pairs == want and 0 <= correct <= pairs
    g.close()
    Xd.free()


@pytest.mark.parametrize("task", [capi.CLASSIFICATION, capi.REGRESSION])
def test_rf_roots_above_the_bucket_table(task):
    """Roots of 2M rows hold more than 4096 x 384 entries: every bucket of the
    chunk pipeline overflows the wave capacity and takes the overfull-bucket
    path; the shallow forest equals the oracle's."""
    if task == capi.REGRESSION:
        X, y = synth.sine_regression(seed=12, n=2_000_000, f=4)
        kw = dict(task=task, n_classes=0)
    else:
        X, y = small_classes(12, 2_000_000, 4, 3)
        kw = dict(n_classes=3)
    g, o = fit_both(X, y, algorithm=capi.RANDOM_FOREST, n_trees=1, k=2, max_depth=2, min_node_size=5, seed=3, **kw)
    assert_same_forest(g, o)
