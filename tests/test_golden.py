"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py):
the digits and iris data files the reference ships, and expected outputs on
them and on seeded synthetic workloads.

CPU tests pin the oracle: on the integer-valued digits every f32 difference,
square and f64 sum is exact, so the distances are closed-form (sqrt of an
integer) and the k kept neighbours are the k nearest; and the oracle still
reproduces every stored output bit for bit.  GPU tests hold the device path
to the same fixtures."""
import os

import numpy as np
import pytest

import oracle as O
from smartcore_amd import synth

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(G, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def digits():
    d = load("digits.npz")
    return d["x"].astype(np.float32), d["y"].astype(np.int32)


@pytest.fixture(scope="module")
def exp():
    return load("expected.npz")


def test_fixture_shapes(digits):
    x, y = digits
    assert x.shape == (1797, 64) and y.shape == (1797,)
    assert set(np.unique(y)) == set(range(10))
    iris = load("iris.npz")
    assert iris["x"].shape == (150, 4) and set(np.unique(iris["y"])) == {0.0, 1.0, 2.0}
    # iris.xy holds the classic measurements (first row 5.1, 3.5, 1.4, 0.2)
    np.testing.assert_array_equal(iris["x"][0], np.float32([5.1, 3.5, 1.4, 0.2]))


def test_digits_distances_are_closed_form(digits, exp):
    x, _ = digits
    q = x[:256].astype(np.int64)
    s = ((q[:, None, :] - x.astype(np.int64)[None, :, :]) ** 2).sum(-1)  # exact integers
    dist = exp["digits_knn_dist"].view(np.float64)
    idx = exp["digits_knn_idx"]
    np.testing.assert_array_equal(dist, np.sqrt(np.take_along_axis(s, idx, 1).astype(np.float64)))
    # the kept set holds the 5 smallest distances (heap order, not sorted)
    np.testing.assert_array_equal(np.sort(dist, 1), np.sqrt(np.sort(s, 1)[:, :5].astype(np.float64)))
    # every query keeps itself (distance 0)
    assert np.all((idx == np.arange(256)[:, None]).any(1) | (dist[:, :1] == 0))
    assert np.all(np.min(dist, 1) == 0.0)


def test_digits_manhattan_closed_form(digits, exp):
    x, _ = digits
    q = x[:64].astype(np.int64)
    s = np.abs(q[:, None, :] - x.astype(np.int64)[None, :, :]).sum(-1)
    dist = exp["digits_l1_dist"].view(np.float64)
    np.testing.assert_array_equal(np.sort(dist, 1), np.sort(s, 1)[:, :7].astype(np.float64))


def test_oracle_reproduces_fixtures(digits, exp):
    x, _ = digits
    i, d, c = O.linear_find_batch(x, x[:256], 5, O.EUCLIDIAN)
    np.testing.assert_array_equal(i, exp["digits_knn_idx"])
    np.testing.assert_array_equal(d.view(np.uint64), exp["digits_knn_dist"])
    i, d, c = O.linear_find_batch(x, x[:64], 7, O.MANHATTAN)
    np.testing.assert_array_equal(i, exp["digits_l1_idx"])
    cent, lab, size, dist, it = O.kmeans_fit(x, x[:10].astype(np.float64), 30, use_bbd=True)
    np.testing.assert_array_equal(lab, exp["digits_km_labels"])
    assert it == int(exp["digits_km_iters"])
    xtr, ytr, xq, yq = synth.knn_mixture(seed=1)
    i, d, c = O.linear_find_batch(xtr, xq[:256], 5, O.EUCLIDIAN)
    np.testing.assert_array_equal(i, exp["c1_idx"])
    np.testing.assert_array_equal(d.view(np.uint64), exp["c1_dist"])
    xk, _, init = synth.kmeans_blobs(seed=5, n=20_000, d=16, blobs=8, k=8)
    np.testing.assert_array_equal(init, exp["km_init"])
    cent, lab, size, dist, it = O.kmeans_fit(xk, init, 20, use_bbd=True)
    np.testing.assert_array_equal(lab, exp["km_labels"])
    np.testing.assert_array_equal(cent.view(np.uint64), exp["km_centroids"].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_device_digits_knn(digits, exp, mode):
    from smartcore_amd import _capi as S
    x, _ = digits
    m = S.SC_MIXED if mode == "mixed" else S.SC_PLAIN_FP64
    ix = S.Index(x)
    i, d, c = ix.knn(x[:256], 5, S.SC_EUCLIDIAN, 2, m)
    np.testing.assert_array_equal(i, exp["digits_knn_idx"])
    np.testing.assert_array_equal(d.view(np.uint64), exp["digits_knn_dist"])
    i, d, c = ix.knn(x[:64], 7, S.SC_MANHATTAN, 1, S.SC_PLAIN_FP64)
    np.testing.assert_array_equal(i, exp["digits_l1_idx"])
    np.testing.assert_array_equal(d.view(np.uint64), exp["digits_l1_dist"])


@pytest.mark.gpu
def test_device_classifier_workload_subset(exp):
    from smartcore_amd import _capi as S
    from smartcore_amd import smartcore as sc
    xtr, ytr, xq, yq = synth.knn_mixture(seed=1)
    ix = S.Index(xtr)
    i, d, c = ix.knn(xq[:256], 5)
    np.testing.assert_array_equal(i, exp["c1_idx"])
    np.testing.assert_array_equal(d.view(np.uint64), exp["c1_dist"])
    p = sc.KNNClassifierParameters().with_k(5).with_algorithm(sc.KNNAlgorithmName.LinearSearch)
    knn = sc.KNNClassifier.fit(xtr, ytr, p)
    np.testing.assert_array_equal(knn.predict(xq[:256]), exp["c1_pred"])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_device_kmeans(digits, exp, mode):
    from smartcore_amd import _capi as S
    m = S.SC_MIXED if mode == "mixed" else S.SC_PLAIN_FP64
    xk, _, init = synth.kmeans_blobs(seed=5, n=20_000, d=16, blobs=8, k=8)
    cent, lab, size, dist, it = S.Index(xk).kmeans_run(init, 20, m)
    np.testing.assert_array_equal(lab, exp["km_labels"])
    np.testing.assert_array_equal(size, exp["km_size"])
    assert it == int(exp["km_iters"])
    ref = exp["km_centroids"]
    assert np.max(np.abs(cent - ref)) <= 1e-12 * np.max(np.abs(ref))
    x, _ = digits
    cent, lab, size, dist, it = S.Index(x).kmeans_run(x[:10].astype(np.float64), 30, m)
    np.testing.assert_array_equal(lab, exp["digits_km_labels"])
    np.testing.assert_array_equal(size, exp["digits_km_size"])
    assert it == int(exp["digits_km_iters"])
    ref = exp["digits_km_centroids"]
    assert np.max(np.abs(cent - ref)) <= 1e-12 * np.max(np.abs(ref))
