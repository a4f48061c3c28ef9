"""Parity of the HIP kNN search (sc_knn_search) with the CPU oracle.

The oracle restates LinearKNNSearch::find (linear_search.rs:52-84) with its
HeapSelection (heap_select.rs); the HIP path must reproduce it bit for bit:
the same indices in the same heap-array order, the same f64 distances and
the same counts.  Minkowski goes through the platform pow on both sides and
is compared on indices with a 1e-12 relative tolerance on distances.
"""
import numpy as np
import pytest

import oracle as O
from smartcore_amd import _capi as S
from smartcore_amd import synth

pytestmark = pytest.mark.gpu

METRIC = {"euclid": (S.SC_EUCLIDIAN, O.EUCLIDIAN), "manhattan": (S.SC_MANHATTAN, O.MANHATTAN),
          "minkowski": (S.SC_MINKOWSKI, O.MINKOWSKI)}


def run_both(data, queries, k, metric="euclid", p=2):
    sm, om = METRIC[metric]
    ix = S.Index(data)
    gi, gd, gc = ix.knn(queries, k, sm, p)
    oi, od, oc = O.linear_find_batch(data, queries, k, om, p)
    ix.close()
    return (gi, gd, gc), (oi, od, oc)


def assert_exact(g, o):
    gi, gd, gc = g
    oi, od, oc = o
    np.testing.assert_array_equal(gc, oc)
    for r in range(gi.shape[0]):
        c = int(oc[r])
        np.testing.assert_array_equal(gi[r, :c], oi[r, :c], err_msg=f"query {r}")
        # bit-exact distances
        np.testing.assert_array_equal(gd[r, :c].view(np.uint64), od[r, :c].view(np.uint64),
                                      err_msg=f"query {r}")


def test_classifier_mixture_f64_k5():
    """KNNClassifier workload: 10,000 x 32 f64, 2,000 queries, k = 5."""
    xtr, ytr, xq, _ = synth.knn_mixture(seed=1)
    ix = S.Index(xtr)
    g = ix.knn(xq, 5)
    st = S.last_stats()
    o = O.linear_find_batch(xtr, xq, 5)
    assert_exact(g, o)
    assert st["queries"] == 2_000 and st["overflow_queries"] == 0
    assert st["candidates"] >= 2_000 * 5


def test_f32_lognormal_k10():
    ref = synth.lognormal_rows(11, 20_000, 128)
    qs = synth.lognormal_rows(12, 128, 128)
    assert_exact(*run_both(ref, qs, 10))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_manhattan(dtype):
    ref = synth.lognormal_rows(21, 8_000, 40).astype(dtype)
    qs = synth.lognormal_rows(22, 100, 40).astype(dtype)
    assert_exact(*run_both(ref, qs, 8, "manhattan"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_minkowski(p, dtype):
    """f32 rows take (a - b) in f32, widened, |.|^p summed in f64
    (minkowski.rs:50-67), like the f64 rows."""
    ref = synth.lognormal_rows(31, 4_000, 24).astype(dtype)
    qs = synth.lognormal_rows(32, 60, 24).astype(dtype)
    (gi, gd, gc), (oi, od, oc) = run_both(ref, qs, 6, "minkowski", p)
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_allclose(gd, od, rtol=1e-12)


def test_tie_heavy_integer_data_keeps_heap_order():
    """Few distinct distances: the kept indices and their order depend on the
    heap's shape, not on index order (e.g. k=5 over 3,0,3,0,2,4,4,4,3,1)."""
    rng = np.random.default_rng(5)
    ref = rng.integers(0, 3, size=(3_000, 4)).astype(np.float64)
    qs = rng.integers(0, 3, size=(150, 4)).astype(np.float64)
    for k in (1, 2, 5, 7, 16):
        assert_exact(*run_both(ref, qs, k))
        assert_exact(*run_both(ref, qs, k, "manhattan"))


def test_tie_counterexample():
    # ten 1-d rows with repeated values: the k = 5 nearest to 0 under
    # Manhattan include a tie at distance 3, which the heap order decides
    ref = np.array([3, 0, 3, 0, 2, 4, 4, 4, 3, 1], np.float64)[:, None]
    ix = S.Index(ref)
    gi, gd, gc = ix.knn(np.zeros((1, 1)), 5, S.SC_MANHATTAN)
    assert sorted(gi[0].tolist()) == [1, 2, 3, 4, 9]
    oi, od = O.linear_find(ref, np.zeros(1), 5, O.MANHATTAN)
    np.testing.assert_array_equal(gi[0], oi)


@pytest.mark.parametrize("d", [1, 3, 16, 17, 33])
def test_odd_dimensions(d):
    ref = synth.lognormal_rows(41 + d, 1_500, d).astype(np.float64)
    qs = synth.lognormal_rows(42 + d, 70, d).astype(np.float64)
    assert_exact(*run_both(ref, qs, 4))


def test_k_edges_and_serial_path():
    ref = synth.lognormal_rows(51, 700, 12).astype(np.float64)
    qs = synth.lognormal_rows(52, 40, 12).astype(np.float64)
    for k in (1, 64, 65, 100, 700):
        assert_exact(*run_both(ref, qs, k))


def test_overflow_falls_back_to_sequential_kernel():
    """Rows at strictly decreasing distance: every row enters the heap, the
    candidate lists overflow and the sequential kernel takes the queries."""
    n = 60_000
    ref = np.arange(n, 0, -1, dtype=np.float64)[:, None]
    # many queries -> few, long segments whose candidate lists overflow
    qs = -np.arange(4_096, dtype=np.float64)[:, None] * 1e-3
    ix = S.Index(ref)
    gi, gd, gc = ix.knn(qs, 10)
    st = S.last_stats()
    oi, od, oc = O.linear_find_batch(ref, qs, 10)
    assert_exact((gi, gd, gc), (oi, od, oc))
    assert st["overflow_queries"] == 4_096


def test_nan_and_inf_rows_are_never_kept():
    ref = np.array([[np.nan, 0.0], [np.inf, 1.0], [1.0, 1.0], [2.0, 2.0], [np.nan, np.nan]])
    qs = np.array([[0.0, 0.0]])
    g, o = run_both(ref, qs, 4)
    assert_exact(g, o)
    assert int(g[2][0]) == 2


def test_find_errors():
    ix = S.Index(np.ones((5, 2)))
    for k in (0, 6):
        with pytest.raises(S.Failed) as e:
            ix.knn(np.ones((1, 2)), k)
        assert e.value.code == S.SC_FIND_FAILED
        assert str(e.value) == "Find failed: k should be >= 1 and <= length(data)"
    with pytest.raises(S.Panic, match="p must be at least 1"):
        ix.knn(np.ones((1, 2)), 2, S.SC_MINKOWSKI, 0)
    with pytest.raises(S.Panic, match="Input vector sizes are different"):
        ix.knn(np.ones((1, 3)), 2)


def test_classifier_vote_on_gpu_neighbours():
    """predict_for_row (knn_classifier.rs:279-299) over the HIP neighbours
    equals the oracle's prediction, for both weight functions."""
    xtr, ytr, xq, _ = synth.knn_mixture(seed=3, n=4_000, q=500)
    classes = np.unique(ytr)
    ycls = np.searchsorted(classes, ytr)
    ix = S.Index(xtr)
    gi, gd, gc = ix.knn(xq, 5)
    oi, od, oc = O.linear_find_batch(xtr, xq, 5)
    for w in (0, 1):
        for r in range(xq.shape[0]):
            a = O.classifier_predict_row(gi[r, :gc[r]], gd[r, :gc[r]], ycls, len(classes), w)
            b = O.classifier_predict_row(oi[r, :oc[r]], od[r, :oc[r]], ycls, len(classes), w)
            assert a == b


@pytest.mark.parametrize("mode,metric", [(S.SC_PLAIN_FP64, S.SC_EUCLIDIAN), (S.SC_MIXED, S.SC_EUCLIDIAN),
                                         (S.SC_MIXED, S.SC_MANHATTAN)])
def test_batch_of_uneven_query_blocks_equals_single_searches(mode, metric):
    """sc_knn_search_batch: blocks of 1..400 rows with k in {1, 5, 10, 20}
    searched in one call give exactly what each gets alone."""
    rng = np.random.default_rng(140)
    ref = (rng.normal(size=(30_000, 40)) * rng.lognormal(0, 0.5, size=(30_000, 1))).astype(np.float32)
    ix = S.Index(ref)
    blocks = []
    for _ in range(25):
        q = int(round(np.exp(rng.uniform(0, np.log(400)))))
        blocks.append((rng.normal(size=(q, 40)).astype(np.float32), int(rng.choice([1, 5, 10, 20]))))
    p = 2 if metric == S.SC_EUCLIDIAN else 1
    got = ix.knn_batch(blocks, metric, p, mode)
    for (qs, k), (gi, gd, gc) in zip(blocks, got):
        ai, ad, ac = ix.knn(qs, k, metric, p, mode)
        np.testing.assert_array_equal(gi, ai)
        np.testing.assert_array_equal(gd.view(np.uint64), ad.view(np.uint64))
        np.testing.assert_array_equal(gc, ac)
    with pytest.raises(S.Failed):
        ix.knn_batch([(blocks[0][0], 5), (blocks[1][0], 0)], metric, p, mode)


def test_device_allocation_failure_is_a_clean_error_and_the_library_recovers():
    """A host-forced device fault: a call whose query buffer (2^40 rows)
    cannot be reserved fails with SC_ERR_OUT_OF_MEMORY before it reads any
    query, and the next calls on the same index and thread work."""
    import ctypes as C

    rng = np.random.default_rng(150)
    ref = rng.normal(size=(2_000, 8))
    qs = rng.normal(size=(3, 8))
    ix = S.Index(ref)
    idx = np.zeros(8, np.int64)
    dist = np.zeros(8)
    cnt = np.zeros(1, np.uint32)
    rc = S.lib().sc_knn_search(ix.handle, qs.ctypes.data_as(C.c_void_p), 1 << 40, 5, S.SC_EUCLIDIAN,
                               2, S.SC_PLAIN_FP64, idx.ctypes.data_as(C.c_void_p),
                               dist.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p))
    assert rc == S.SC_ERR_OUT_OF_MEMORY
    assert "HIP error" in S.last_error()
    for mode in (S.SC_PLAIN_FP64, S.SC_MIXED):
        gi, gd, gc = ix.knn(qs, 5, S.SC_EUCLIDIAN, 2, mode)
        oi, od, oc = O.linear_find_batch(ref, qs, 5)
        np.testing.assert_array_equal(gi, oi)


def test_concurrent_searches_on_one_index_from_several_threads():
    """Calls on one index from several host threads run on per-thread streams
    with per-call workspaces: every answer equals the same search run alone.
    The index is fresh, so the first mixed calls also race to build its BF16
    copy (built once, under the index's lock)."""
    import threading

    ref = synth.lognormal_rows(61, 60_000, 48)
    qs = synth.lognormal_rows(62, 2_000, 48)
    jobs = [(S.SC_EUCLIDIAN, S.SC_MIXED, 10), (S.SC_EUCLIDIAN, S.SC_PLAIN_FP64, 7),
            (S.SC_MANHATTAN, S.SC_MIXED, 5), (S.SC_MANHATTAN, S.SC_PLAIN_FP64, 3)]
    ix0 = S.Index(ref)
    want = {}
    for j, (m, mode, k) in enumerate(jobs):
        want[j] = ix0.knn(qs[200 * j:200 * j + 600], k, m, 2, mode)
    ix0.close()

    ix = S.Index(ref)
    got = {}
    errors = []

    def worker(t):
        try:
            for rep in range(3):
                j = (t + rep) % len(jobs)
                m, mode, k = jobs[j]
                got[(t, rep)] = (j, ix.knn(qs[200 * j:200 * j + 600], k, m, 2, mode))
        except Exception as e:  # reported below, with the thread
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    ix.close()
    assert not errors, errors
    assert len(got) == 18
    for (t, rep), (j, (gi, gd, gc)) in got.items():
        wi, wd, wc = want[j]
        np.testing.assert_array_equal(gc, wc, err_msg=f"thread {t} call {rep}")
        np.testing.assert_array_equal(gi, wi, err_msg=f"thread {t} call {rep}")
        np.testing.assert_array_equal(gd.view(np.uint64), wd.view(np.uint64),
                                      err_msg=f"thread {t} call {rep}")
