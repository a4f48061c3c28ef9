"""SC_MIXED k-means (certified BF16 MFMA assignment fused with the FP64 sums)
must reproduce the FP64 form bit for bit -- labels, sizes, iteration count,
centroid bits, distortion bits -- and the BBD-tree oracle's labels."""
import numpy as np
import pytest

import oracle as O
from smartcore_amd import _capi as S
from smartcore_amd import synth

pytestmark = pytest.mark.gpu


def run_both(x, init, iters):
    ix = S.Index(x)
    a = ix.kmeans_run(init, iters, S.SC_PLAIN_FP64)
    b = ix.kmeans_run(init, iters, S.SC_MIXED)
    st = S.last_stats()
    return a, b, st


def assert_same(a, b):
    ca, la, sa, da, ia = a
    cb, lb, sb, db, ib = b
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(sa, sb)
    assert ia == ib
    np.testing.assert_array_equal(ca.view(np.uint64), cb.view(np.uint64))
    assert np.float64(da).view(np.uint64) == np.float64(db).view(np.uint64)


@pytest.mark.parametrize("n,d,k", [(50_000, 64, 64), (20_000, 16, 12), (30_000, 24, 70),
                                   (9_000, 100, 5), (4_097, 8, 128), (3_000, 100, 100)])
def test_mixed_equals_fp64_form(n, d, k):
    x, _, init = synth.kmeans_blobs(seed=20 + k, n=n, d=d, blobs=max(2, k // 2), k=k)
    a, b, st = run_both(x, init, 15)
    assert_same(a, b)


def test_mixed_labels_match_bbd_oracle():
    x, _, init = synth.kmeans_blobs(seed=31, n=40_000, d=32, blobs=16, k=16)
    ix = S.Index(x)
    cent, labels, size, dist, iters = ix.kmeans_run(init, 20, S.SC_MIXED)
    ocent, olabels, osize, odist, oiters = O.kmeans_fit(x, init, 20, use_bbd=True)
    np.testing.assert_array_equal(labels, olabels)
    np.testing.assert_array_equal(size, osize)
    assert iters == oiters
    assert np.max(np.abs(cent - ocent)) <= 1e-12 * np.max(np.abs(ocent))


@pytest.mark.parametrize("mode", [S.SC_MIXED, S.SC_PLAIN_FP64])
def test_config2_shape_vs_bbd_oracle(mode):
    """The 1,000,000 x 64, k = 64 workload's shape class (d = 64, k = 64,
    64 blobs of U(0.5, 2) spread) on 60,000 rows, 20 Lloyd iterations from the
    same initial centroids: labels, sizes and iteration count equal the
    BBD-tree oracle's; the centroids agree to 1e-12 relative (the oracle adds
    the sums in tree order, the device in fixed row-block order)."""
    x, _, init = synth.kmeans_blobs(seed=2, n=60_000, d=64, blobs=64, k=64)
    ix = S.Index(x)
    cent, labels, size, dist, iters = ix.kmeans_run(init, 20, mode)
    ocent, olabels, osize, odist, oiters = O.kmeans_fit(x, init, 20, use_bbd=True)
    np.testing.assert_array_equal(labels, olabels)
    np.testing.assert_array_equal(size, osize)
    assert iters == oiters
    rel = np.max(np.abs(cent - ocent)) / np.max(np.abs(ocent))
    print(f"centroid relative error vs oracle: {rel:.3e}; distortion {dist!r} vs {odist!r}")
    assert rel <= 1e-12
    assert abs(dist - odist) <= 1e-9 * abs(odist)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_stop_rule_on_data_far_from_the_origin(dtype):
    """Blobs shifted by 1e5 in every dimension, up to 200 iterations: the
    distortion compared by the stop rule (kmeans.rs:305-309) must not lose
    the step-to-step decrease to cancellation, so the fit stops at the
    oracle's iteration with the oracle's labels (160 for f64 rows, 132 for
    f32 rows, both before max_iter)."""
    x, _, init = synth.kmeans_blobs(seed=21, n=30_000, d=16, blobs=10, k=10, dtype=np.float64)
    x = (x + 1e5).astype(dtype)
    init = init + 1e5
    ocent, olabels, osize, odist, oiters = O.kmeans_fit(x, init, 200, use_bbd=True)
    assert oiters < 200
    ix = S.Index(x)
    modes = [S.SC_PLAIN_FP64] + ([S.SC_MIXED] if dtype == np.float32 else [])
    for mode in modes:
        cent, labels, size, dist, iters = ix.kmeans_run(init, 200, mode)
        assert iters == oiters
        np.testing.assert_array_equal(labels, olabels)
        np.testing.assert_array_equal(size, osize)
        assert abs(dist - odist) <= 1e-9 * abs(odist)


def test_mixed_near_ties_use_fp64_fallback():
    """Centroids placed almost on top of each other cannot be separated by
    the BF16 bound, so those rows go through the FP64 path."""
    rng = np.random.default_rng(40)
    x = rng.normal(size=(20_000, 32)).astype(np.float32)
    base = x[:8].astype(np.float64)
    init = np.concatenate([base, base + 1e-9]).astype(np.float64)
    a, b, st = run_both(x, init, 5)
    assert_same(a, b)
    assert st["uncertified"] > 0


def test_mixed_exact_ties_take_lowest_index():
    """Duplicated centroids tie exactly for every row that prefers them, so
    neither bound separates them and the sequential FP64 fold decides: the
    lowest index must win, as in the FP64 form."""
    rng = np.random.default_rng(43)
    x = rng.normal(size=(12_000, 24)).astype(np.float32)
    base = x[:6].astype(np.float64)
    init = np.concatenate([base, base[:3]]).astype(np.float64)
    a, b, st = run_both(x, init, 6)
    assert_same(a, b)
    assert st["uncertified"] > 0


def test_mixed_handles_non_finite_rows():
    rng = np.random.default_rng(41)
    x = rng.normal(size=(5_000, 16)).astype(np.float32)
    x[7, 3] = np.nan
    x[100, 0] = np.inf
    init = x[200:204].astype(np.float64)
    a, b, st = run_both(x, init, 3)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])


def test_mixed_on_f64_rows_serves_fp64_form():
    x, _, init = synth.kmeans_blobs(seed=42, n=10_000, d=16, blobs=8, k=8, dtype=np.float64)
    a, b, st = run_both(x, init, 10)
    assert_same(a, b)


def test_mixed_tiny_and_mixed_scale_rows():
    """Rows with squared norms below 1e-30 (whose BF16 split terms could
    underflow) take the all-centroid path; rows of ordinary scale in the same
    set keep the certified path.  Both must match the FP64 form."""
    rng = np.random.default_rng(44)
    x = rng.normal(size=(8_000, 32)).astype(np.float32)
    x[:2_000] *= np.float32(1e-17)   # ||x||^2 ~ 3e-32
    x[2_000:2_100] = 0.0             # exact zeros stay certified
    init = np.concatenate([x[:3], x[4_000:9_000:1_000]]).astype(np.float64)
    a, b, st = run_both(x, init, 6)
    assert_same(a, b)
