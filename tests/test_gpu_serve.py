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
