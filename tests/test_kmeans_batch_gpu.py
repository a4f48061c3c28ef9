"""Batched k-means (sc_kmeans_batch_run): many uneven problems in one
persistent kernel per iteration, captured in a hipGraph.  Every problem's
result must equal the single-problem SC_MIXED fit bit for bit, and its labels
the FP64 form's."""
import numpy as np
import pytest

import oracle as O
from smartcore_amd import _capi as S
from smartcore_amd import synth

pytestmark = pytest.mark.gpu


def single(x, init, iters, mode):
    ix = S.Index(x)
    return ix.kmeans_run(init, iters, mode)


def test_batch_equals_single_problem_runs():
    probs = synth.kmeans_batch_problems(seed=50, count=24, n_lo=500, n_hi=40_000)
    got = S.kmeans_batch_run(probs, 8, S.SC_MIXED)
    for (x, init), (c, lab, size, dist, it) in zip(probs, got):
        c1, l1, s1, d1, i1 = single(x, init, 8, S.SC_MIXED)
        np.testing.assert_array_equal(lab, l1)
        np.testing.assert_array_equal(size, s1)
        assert it == i1
        np.testing.assert_array_equal(c.view(np.uint64), c1.view(np.uint64))
        assert np.float64(dist).view(np.uint64) == np.float64(d1).view(np.uint64)
        c2, l2, s2, d2, i2 = single(x, init, 8, S.SC_PLAIN_FP64)
        np.testing.assert_array_equal(lab, l2)


def test_batch_second_call_uses_tuned_grid_and_agrees():
    probs = synth.kmeans_batch_problems(seed=51, count=10, n_lo=300, n_hi=9_000)
    a = S.kmeans_batch_run(probs, 6, S.SC_MIXED)
    b = S.kmeans_batch_run(probs, 6, S.SC_MIXED)
    for (ca, la, sa, da, ia), (cb, lb, sb, db, ib) in zip(a, b):
        np.testing.assert_array_equal(la, lb)
        np.testing.assert_array_equal(ca.view(np.uint64), cb.view(np.uint64))
        assert ia == ib


def test_batch_launch_shape_tuner_is_invariant_and_reported():
    """The first call on a new set of shape classes times six label grid
    sizes (1, 2, 3, 4, 6 or 8 four-wave workgroups per CU) on its first six
    iterations, keeps the fastest and reports it; every problem must still
    equal its single-problem fit (which runs as many workgroups as a CU
    holds), and a second call (the kept shape for all iterations) must
    repeat the first bit for bit.  The problems cover every k block count
    and the widest rows, with k and d up to 128 and 100."""
    rng = np.random.default_rng(53)
    probs = []
    for n, d, k in [(5_000, 8, 3), (12_000, 24, 40), (9_000, 64, 64), (7_000, 40, 100),
                    (6_000, 100, 20), (3_000, 64, 128), (20_000, 16, 12), (4_000, 100, 100)]:
        x = rng.normal(size=(n, d)).astype(np.float32)
        x += rng.integers(0, 5, size=(n, 1)).astype(np.float32) * 3.0
        init = x[rng.choice(n, size=k, replace=False)].astype(np.float64)
        probs.append((x, init))
    a = S.kmeans_batch_run(probs, 9, S.SC_MIXED)
    st = S.last_stats()
    assert st["launch_waves"] == 4
    assert st["launch_groups_per_cu"] in (1, 2, 3, 4, 6, 8)
    b = S.kmeans_batch_run(probs, 9, S.SC_MIXED)
    st2 = S.last_stats()
    assert (st2["launch_waves"], st2["launch_groups_per_cu"]) == (st["launch_waves"],
                                                                 st["launch_groups_per_cu"])
    for (x, init), ra, rb in zip(probs, a, b):
        c1, l1, s1, d1, i1 = single(x, init, 9, S.SC_MIXED)
        for c, lab, size, dist, it in (ra, rb):
            np.testing.assert_array_equal(lab, l1)
            np.testing.assert_array_equal(size, s1)
            assert it == i1
            np.testing.assert_array_equal(c.view(np.uint64), c1.view(np.uint64))
            assert np.float64(dist).view(np.uint64) == np.float64(d1).view(np.uint64)


def test_large_batch_equals_single_runs():
    """Six problems of 200,000 x 64 f32 rows (307 MB, several hundred
    label items per problem, rows staged through the pinned buffers in
    many pieces, in two groups: the second group's rows upload while the
    first group iterates): every problem equals its single-problem fit bit for bit
    and its labels the FP64 form's, both on the first call (six tuning
    iterations per group, then a graph) and on a second call (the kept
    shapes, every iteration in the graph)."""
    rng = np.random.default_rng(55)
    probs = []
    for j in range(6):
        n, d, k = 200_000, 64, 16 + 8 * j
        centres = rng.normal(0.0, 4.0, size=(k, d))
        x = (centres[rng.integers(0, k, size=n)] + rng.normal(size=(n, d))).astype(np.float32)
        init = x[np.sort(rng.choice(n, size=k, replace=False))].astype(np.float64)
        probs.append((x, init))
    assert sum(x.nbytes for x, _ in probs) > 256 << 20
    iters = 9
    first = S.kmeans_batch_run(probs, iters, S.SC_MIXED)
    st = S.last_stats()
    assert st["queries"] == len(probs)
    assert st["upload_groups"] == 2
    second = S.kmeans_batch_run(probs, iters, S.SC_MIXED)
    assert S.last_stats()["upload_groups"] == 2
    for (x, init), ra, rb in zip(probs, first, second):
        c1, l1, s1, d1, i1 = single(x, init, iters, S.SC_MIXED)
        for c, lab, size, dist, it in (ra, rb):
            np.testing.assert_array_equal(lab, l1)
            np.testing.assert_array_equal(size, s1)
            assert it == i1
            np.testing.assert_array_equal(c.view(np.uint64), c1.view(np.uint64))
            assert np.float64(dist).view(np.uint64) == np.float64(d1).view(np.uint64)
        c2, l2, s2, d2, i2 = single(x, init, iters, S.SC_PLAIN_FP64)
        np.testing.assert_array_equal(ra[1], l2)


def test_batch_group_of_one_problem():
    """One problem of 500,000 x 64 rows (128 MB, nine tenths of the bytes)
    and three small ones: the first group holds the large problem alone, so
    its reduce forms the group sums inside the final kernel, as a single
    fit does; the second group's three problems take the per-problem close
    kernel.  Every problem equals its single-problem fit bit for bit."""
    rng = np.random.default_rng(57)
    probs = []
    for n, k in [(500_000, 48), (20_000, 16), (15_000, 100), (25_000, 7)]:
        d = 64
        centres = rng.normal(0.0, 4.0, size=(k, d))
        x = (centres[rng.integers(0, k, size=n)] + rng.normal(size=(n, d))).astype(np.float32)
        init = x[np.sort(rng.choice(n, size=k, replace=False))].astype(np.float64)
        probs.append((x, init))
    iters = 8
    got = S.kmeans_batch_run(probs, iters, S.SC_MIXED)
    assert S.last_stats()["upload_groups"] == 2
    for (x, init), (c, lab, size, dist, it) in zip(probs, got):
        c1, l1, s1, d1, i1 = single(x, init, iters, S.SC_MIXED)
        np.testing.assert_array_equal(lab, l1)
        np.testing.assert_array_equal(size, s1)
        assert it == i1
        np.testing.assert_array_equal(c.view(np.uint64), c1.view(np.uint64))
        assert np.float64(dist).view(np.uint64) == np.float64(d1).view(np.uint64)


def test_batch_fp64_mode_runs_problems_one_by_one():
    probs = synth.kmeans_batch_problems(seed=52, count=4, n_lo=300, n_hi=3_000)
    got = S.kmeans_batch_run(probs, 5, S.SC_PLAIN_FP64)
    for (x, init), (c, lab, size, dist, it) in zip(probs, got):
        c1, l1, s1, d1, i1 = single(x, init, 5, S.SC_PLAIN_FP64)
        np.testing.assert_array_equal(lab, l1)
        np.testing.assert_array_equal(c.view(np.uint64), c1.view(np.uint64))


def test_batch_errors():
    x = np.ones((10, 4), np.float32)
    with pytest.raises(S.Failed):
        S.kmeans_batch_run([(x, np.ones((1, 4)))], 5)
    with pytest.raises(S.Failed):
        S.kmeans_batch_run([(x, np.ones((2, 4)))], 0)


def test_batch_with_a_problem_outside_the_shape_classes():
    # k = 130 is above the 128 clusters the MFMA form handles: the batch runs
    # its problems one by one, each equal to its single run
    probs = synth.kmeans_batch_problems(seed=53, count=3, n_lo=300, n_hi=2_000)
    probs += synth.kmeans_batch_problems(seed=54, count=1, n_lo=1_500, n_hi=1_500, dims=(100,),
                                         k_lo=130, k_hi=130)
    got = S.kmeans_batch_run(probs, 4, S.SC_MIXED)
    for (x, init), (c, lab, size, dist, it) in zip(probs, got):
        c1, l1, s1, d1, i1 = single(x, init, 4, S.SC_MIXED)
        np.testing.assert_array_equal(lab, l1)
        np.testing.assert_array_equal(c.view(np.uint64), c1.view(np.uint64))
        assert it == i1
        c2, l2, s2, d2, i2 = single(x, init, 4, S.SC_PLAIN_FP64)
        np.testing.assert_array_equal(lab, l2)


def test_batch_config4_problems_vs_bbd_oracle():
    """Sixteen problems drawn like the 512-problem workload (n log-uniform in
    [1,000, 200,000], d in {8, 16, 32, 64}, k in 4..128) plus three with n
    near 200,000, 20 iterations in one batched call: every problem's labels,
    sizes and iteration count equal the BBD-tree oracle's fit, centroids to
    1e-12 relative."""
    probs = synth.kmeans_batch_problems(seed=60, count=16)
    probs += synth.kmeans_batch_problems(seed=61, count=3, n_lo=180_000, n_hi=200_000)
    assert {x.shape[1] for x, _ in probs} == {8, 16, 32, 64}
    got = S.kmeans_batch_run(probs, 20, S.SC_MIXED)
    worst = 0.0
    for (x, init), (c, lab, size, dist, it) in zip(probs, got):
        ocent, olabels, osize, odist, oiters = O.kmeans_fit(x, init, 20, use_bbd=True)
        np.testing.assert_array_equal(lab, olabels)
        np.testing.assert_array_equal(size, osize)
        assert it == oiters
        rel = np.max(np.abs(c - ocent)) / np.max(np.abs(ocent))
        worst = max(worst, rel)
        assert rel <= 1e-12
    print(f"worst centroid relative error vs oracle over {len(probs)} problems: {worst:.3e}")
