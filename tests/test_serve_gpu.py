"""Continuous-batching kNN server: every request's answer equals the same
search run alone (sc_knn_search), whatever else shares its batch; admission
by HBM budget queues requests instead of failing them; an oversized request
is refused."""
import threading

import numpy as np
import pytest

import oracle as O
from smartcore_amd import _capi as S

pytestmark = pytest.mark.gpu


def make(n=60_000, d=96, seed=90):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, d)).astype(np.float32), rng


def test_results_same_alone_and_in_the_mix():
    ref, rng = make()
    ix = S.Index(ref)
    srv = S.Server(ix, max_batch_queries=2048, max_inflight=3)
    reqs = []
    for i in range(40):
        q = int(round(np.exp(rng.uniform(0, np.log(600)))))
        metric = S.SC_MANHATTAN if i % 5 == 0 else S.SC_EUCLIDIAN
        # Manhattan requests alternate between the bounded f32 form and the
        # FP64 form
        mode = S.SC_PLAIN_FP64 if metric == S.SC_MANHATTAN and i % 10 == 0 else S.SC_MIXED
        reqs.append((rng.normal(size=(q, 96)).astype(np.float32), metric, mode))
    results = {}

    def client(part):
        tickets = [(j, srv.submit(reqs[j][0], 20, reqs[j][1], 1 if reqs[j][1] else 2, reqs[j][2]))
                   for j in part]
        for j, t in tickets:
            results[j] = srv.wait(t)

    threads = [threading.Thread(target=client, args=(list(range(c, 40, 4)),)) for c in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    st = srv.stats()
    srv.close()
    assert st["completed"] == 40
    assert st["batches"] < 40  # requests were coalesced
    for j, (qs, metric, mode) in enumerate(reqs):
        gi, gd, gc, lat = results[j]
        ai, ad, ac = ix.knn(qs, 20, metric, 1 if metric else 2, mode)
        np.testing.assert_array_equal(gi, ai)
        np.testing.assert_array_equal(gd.view(np.uint64), ad.view(np.uint64))
        np.testing.assert_array_equal(gc, ac)
        assert lat > 0.0
    # and the oracle on a few rows of a Euclidean and a mixed Manhattan
    # request from the mix
    oi, od, oc = O.linear_find_batch(ref, reqs[1][0][:3], 20)
    np.testing.assert_array_equal(results[1][0][:3], oi)
    assert reqs[5][1] == S.SC_MANHATTAN and reqs[5][2] == S.SC_MIXED
    oi, od, oc = O.linear_find_batch(ref, reqs[5][0][:3], 20, O.MANHATTAN)
    np.testing.assert_array_equal(results[5][0][:3], oi)
    np.testing.assert_array_equal(results[5][1][:3].view(np.uint64), od.view(np.uint64))


def test_admission_queues_under_a_small_budget():
    ref, rng = make(n=20_000, d=32, seed=91)
    ix = S.Index(ref)
    reqs = [rng.normal(size=(256, 32)).astype(np.float32) for _ in range(12)]
    # each request needs ~12 MB of workspace by the server's estimate: a
    # 20 MB budget admits one at a time although three slots are free
    srv = S.Server(ix, max_batch_queries=256, max_inflight=3, hbm_budget_bytes=20 << 20)
    tickets = [srv.submit(r, 10) for r in reqs]
    outs = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 12 and st["refused"] == 0
    assert st["hbm_peak"] <= st["hbm_budget"]
    for r, (gi, gd, gc, lat) in zip(reqs, outs):
        ai, ad, ac = ix.knn(r, 10)
        np.testing.assert_array_equal(gi, ai)


def test_oversized_request_is_refused():
    ref, rng = make(n=5_000, d=16, seed=92)
    ix = S.Index(ref)
    srv = S.Server(ix, max_batch_queries=4096, hbm_budget_bytes=1 << 16)
    with pytest.raises(S.DeviceError):
        srv.submit(rng.normal(size=(4000, 16)).astype(np.float32), 10)
    assert srv.stats()["refused"] == 1
    with pytest.raises(S.Failed):
        srv.submit(rng.normal(size=(4, 16)).astype(np.float32), 0)
    srv.close()
