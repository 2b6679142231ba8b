"""Serving: resident forests, continuous batching of uneven requests.

Every result must equal the forest's own predict on the same rows, bit for
bit, whether the request ran alone or in a mix of requests for several
forests submitted from several threads while other batches were on the GPU.
"""
import threading

import numpy as np
import pytest

import pyoracle
from datamininglibs_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def forests():
    Xc, yc = synth.gaussian_classes(seed=21, n=4000, f=24, n_informative=8, n_classes=5)
    Xr, yr = synth.sine_regression(seed=22, n=3000, f=24)
    fc = capi.fit(Xc, yc, capi.default_params(capi.RANDOM_FOREST, n_trees=40, n_features_per_node=5, max_depth=12,
                                                n_classes=5, seed=3))
    fe = capi.fit(Xc, yc, capi.default_params(capi.EXTRA_TREES, n_trees=60, n_classes=5, seed=4))
    fr = capi.fit(Xr, yr, capi.default_params(capi.RANDOM_FOREST, task=capi.REGRESSION, n_trees=30,
                                                n_features_per_node=5, max_depth=10, n_classes=0, seed=5))
    return {"cls_rf": fc, "cls_ert": fe, "reg_rf": fr}


def rows_for(rng, n, f=24):
    return rng.standard_normal((n, f)).astype(np.float32)


def alone(forest, rows):
    out = forest.predict(np.ascontiguousarray(rows.T))
    return out


def test_single_request_matches_predict(forests):
    srv = capi.Server(max_batch_rows=1024)
    try:
        rng = np.random.default_rng(1)
        for name, f in forests.items():
            slot = srv.load(f)
            for n in (1, 7, 1024, 1025, 3000):
                rows = rows_for(rng, n)
                got = srv.wait(srv.submit(slot, rows))
                assert np.array_equal(got, alone(f, rows)), (name, n)
        st = srv.stats()
        assert st["requests_done"] == 15 and st["rows_done"] == 3 * (1 + 7 + 1024 + 1025 + 3000)
        assert st["hbm_in_use"] > 0
    finally:
        srv.close()


def test_mixed_concurrent_requests_equal_alone(forests):
    srv = capi.Server(max_batch_rows=2048)
    names = list(forests)
    slots = {n: srv.load(forests[n]) for n in names}
    rng = np.random.default_rng(7)
    work = []
    for i in range(120):
        name = names[int(rng.integers(0, len(names)))]
        n = int(np.round(np.exp(rng.uniform(0, np.log(6000)))))
        work.append((name, rows_for(rng, max(1, n))))
    results = [None] * len(work)
    errors = []

    def client(idx):
        try:
            tickets = []
            for i in idx:
                name, rows = work[i]
                tickets.append((i, srv.submit(slots[name], rows)))
            for j, (i, t) in enumerate(tickets):
                if j % 2:
                    results[i] = srv.wait(t)
                else:
                    while True:
                        r = srv.poll(t)
                        if r is not None:
                            results[i] = r
                            break
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=client, args=(list(range(k, len(work), 4)),)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    st = srv.stats()
    srv.close()
    for i, (name, rows) in enumerate(work):
        assert np.array_equal(results[i], alone(forests[name], rows)), (i, name, rows.shape)
    assert st["requests_done"] == len(work)
    assert st["batches_launched"] < len(work) + sum(r.shape[0] for _, r in work) // 2048 + 1


def test_server_results_equal_oracle_votes():
    X, y = synth.spam_like(seed=0, n=1500, f=16)
    f = capi.fit(X, y, capi.default_params(capi.EXTRA_TREES, n_trees=20, n_classes=2, seed=9))
    o = pyoracle.fit(X, y, algorithm=1, n_trees=20, n_classes=2, seed=9)
    srv = capi.Server(max_batch_rows=256)
    try:
        slot = srv.load(f)
        Xp, _ = synth.spam_like(seed=2, n=700, f=16)
        got = srv.wait(srv.submit(slot, np.ascontiguousarray(Xp.T)))
        assert np.array_equal(got.astype(np.int64), o.predict(Xp).astype(np.int64))
    finally:
        srv.close()


def test_admission_by_budget(forests):
    srv = capi.Server(hbm_budget_bytes=1 << 16, max_batch_rows=16)
    try:
        with pytest.raises(capi.DteError) as e:
            srv.load(forests["cls_ert"])
        assert e.value.code == capi.ERR_ADMISSION
        assert srv.stats()["requests_rejected"] == 1
    finally:
        srv.close()


def test_unknown_ticket_and_slot(forests):
    srv = capi.Server(max_batch_rows=64)
    try:
        with pytest.raises(capi.DteError) as e:
            srv.submit(3, np.zeros((2, 24), np.float32))
        assert e.value.code == capi.ERR_NOT_FOUND
        slot = srv.load(forests["reg_rf"])
        req = srv.submit(slot, np.zeros((2, 24), np.float32))
        srv.wait(req)
        with pytest.raises(capi.DteError) as e:
            srv.wait(req)  # already delivered
        assert e.value.code == capi.ERR_NOT_FOUND
    finally:
        srv.close()


def test_load_forest_while_requests_in_flight(forests):
    """Forests loaded while a client keeps batches on the GPU: the server's
    forest table grows under the scheduler, every result still equals the
    forest's own predict."""
    srv = capi.Server(max_batch_rows=512)
    try:
        first = srv.load(forests["cls_rf"])
        rng = np.random.default_rng(11)
        work = [rows_for(rng, int(rng.integers(1, 3000))) for _ in range(60)]
        results, errors = [None] * len(work), []

        def client():
            try:
                tickets = [srv.submit(first, rows) for rows in work]
                for i, t in enumerate(tickets):
                    results[i] = srv.wait(t)
            except Exception as e:  # surfaced below
                errors.append(e)

        th = threading.Thread(target=client)
        th.start()
        extra = []
        for _ in range(6):  # the table grows past its first capacity
            for name in ("cls_ert", "reg_rf"):
                extra.append((name, srv.load(forests[name])))
        th.join(timeout=120)
        assert not errors, errors
        for i, rows in enumerate(work):
            assert np.array_equal(results[i], alone(forests["cls_rf"], rows)), i
        rows = rows_for(rng, 700)
        for name, slot in extra[-2:]:
            assert np.array_equal(srv.wait(srv.submit(slot, rows)), alone(forests[name], rows)), name
    finally:
        srv.close()


def test_batches_carry_several_forests(forests):
    """Small requests for three forests submitted before the scheduler can
    drain them ride together: a batch launches one traversal per forest it
    carries, and every result equals the forest's own predict."""
    srv = capi.Server(max_batch_rows=4096)
    try:
        names = list(forests)
        slots = {n: srv.load(forests[n]) for n in names}
        rng = np.random.default_rng(12)
        work = [(names[i % 3], rows_for(rng, int(rng.integers(1, 40)))) for i in range(300)]
        tickets = [srv.submit(slots[n], rows) for n, rows in work]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
        for (n, rows), g in zip(work, got):
            assert np.array_equal(g, alone(forests[n], rows)), n
        assert st["groups_launched"] > st["batches_launched"]
        assert st["kernel_ms"] > 0 and st["h2d_ms"] > 0 and st["d2h_ms"] > 0
        assert abs(st["busy_ms"] - (st["h2d_ms"] + st["kernel_ms"] + st["d2h_ms"])) < 1e-6 * max(1.0, st["busy_ms"])
    finally:
        srv.close()


def test_two_waiters_on_one_ticket(forests):
    """Two callers wait on the same ticket: one receives the result, the
    other a not-found error; nothing reads a request after it was freed."""
    srv = capi.Server(max_batch_rows=1024)
    try:
        slot = srv.load(forests["cls_ert"])
        rng = np.random.default_rng(13)
        for _ in range(20):
            rows = rows_for(rng, 5000)
            req = srv.submit(slot, rows)
            outcomes = []

            def waiter():
                try:
                    outcomes.append(("ok", srv.wait(req)))
                except capi.DteError as e:
                    outcomes.append(("err", e.code))

            ths = [threading.Thread(target=waiter) for _ in range(2)]
            for t in ths:
                t.start()
            for t in ths:
                t.join(timeout=60)
            kinds = sorted(k for k, _ in outcomes)
            assert kinds == ["err", "ok"], outcomes
            ok = next(v for k, v in outcomes if k == "ok")
            err = next(v for k, v in outcomes if k == "err")
            assert err == capi.ERR_NOT_FOUND
            assert np.array_equal(ok, alone(forests["cls_ert"], rows))
    finally:
        srv.close()


def test_serving_shape_500_tree_forests_256_features():
    """The serving workload's shape: 4 resident forests of 500 trees (depth
    <= 20, 256 features, 10 classes) admitted under an explicit HBM budget;
    requests of log-uniform 1..65,536 rows (one of exactly 65,536) from four
    threads.  Every result equals its forest's own predict; forest 0 equals
    the oracle's forest and a row subset of its served votes equals the
    oracle's votes (gpudte_algorithm.cc:269-331, gpudte.cu:311-359)."""
    from parity import assert_same_forest
    n_train = 20_000
    X, y = synth.gaussian_classes(seed=5, n=n_train, f=256, n_informative=16, n_classes=10)
    forests = [capi.fit(X, y, capi.default_params(capi.RANDOM_FOREST, n_trees=500, max_depth=20, min_node_size=5,
                                                  n_classes=10, seed=1000 + i)) for i in range(4)]
    o = pyoracle.fit(X, y, algorithm=0, n_trees=500, max_depth=20, min_node_size=5, n_classes=10, seed=1000,
                     n_threads=16)
    assert_same_forest(forests[0], o)
    need = sum(f.n_nodes * 16 + 8 * 501 for f in forests) + 4 * 65536 * (256 + 10) * 4
    srv = capi.Server(hbm_budget_bytes=need + (64 << 20), max_batch_rows=65536)
    try:
        slots = [srv.load(f) for f in forests]
        assert srv.stats()["hbm_in_use"] <= need + (64 << 20)
        rng = np.random.default_rng(14)
        sizes = list(synth.request_sizes(seed=15, n_requests=48, max_rows=65536)) + [65536]
        work = [(int(rng.integers(0, 4)), rng.standard_normal((int(n), 256), dtype=np.float32)) for n in sizes]
        results, errors = [None] * len(work), []

        def client(idx):
            try:
                tickets = [(i, srv.submit(slots[work[i][0]], work[i][1])) for i in idx]
                for i, t in tickets:
                    results[i] = srv.wait(t)
            except Exception as e:  # surfaced below
                errors.append(e)

        ths = [threading.Thread(target=client, args=(list(range(k, len(work), 4)),)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(timeout=300)
        assert not errors, errors
        for i, (fi, rows) in enumerate(work):
            assert np.array_equal(results[i], alone(forests[fi], rows)), (i, fi, rows.shape)
            assert (results[i].sum(1) == 500).all()
        sub = [i for i, (fi, _) in enumerate(work) if fi == 0][:6]
        for i in sub:
            rows = work[i][1][:2000]
            want = o.predict(np.ascontiguousarray(rows.T)).astype(np.int64)
            assert np.array_equal(results[i][:2000].astype(np.int64), want), i
        assert srv.stats()["requests_done"] == len(work)
    finally:
        srv.close()
        for f in forests:
            f.close()
