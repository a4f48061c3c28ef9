"""DBSCAN (cluster/dbscan.rs:274-397) on the device radius search: labels,
cluster count and predictions equal the oracle's restatement of the
reference loop; the reference's own test through the C++ mirror."""
import os

import numpy as np
import pytest

import oracle as O
from smartcore_amd import _capi as S
from smartcore_amd import smartcore as sc
from smartcore_amd import synth

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_fit_predict_dbscan_kat():
    """dbscan.rs:430-470"""
    x = np.array([[1.0, 2.0], [1.1, 2.1], [0.9, 1.9], [1.2, 2.2], [0.8, 1.8], [2.0, 1.0],
                  [2.1, 1.1], [1.9, 0.9], [2.2, 1.2], [1.8, 0.8], [3.0, 5.0]])
    p = sc.DBSCANParameters().with_eps(0.5).with_min_samples(2)
    m = sc.DBSCAN.fit(x, p)
    np.testing.assert_array_equal(m.predict(x), [1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 0])
    assert m.num_classes == 2


def test_dbscan_errors():
    x = np.ones((4, 2))
    with pytest.raises(sc.Failed, match="Invalid minPts"):
        sc.DBSCAN.fit(x, sc.DBSCANParameters().with_min_samples(0))
    with pytest.raises(sc.Failed, match="Invalid radius"):
        sc.DBSCAN.fit(x, sc.DBSCANParameters().with_eps(0.0))


@pytest.mark.parametrize("eps,ms", [(18.0, 5), (22.0, 12), (25.0, 30)])
def test_digits_equals_oracle(eps, ms):
    x = np.load(os.path.join(G, "digits.npz"))["x"].astype(np.float32)
    fit, q = x[:1200], x[1200:]
    ix = S.Index(fit)
    lab, nc = ix.dbscan_fit(eps, ms)
    olab, onc = O.dbscan_fit(fit, eps, ms)
    assert nc == onc
    np.testing.assert_array_equal(lab, olab)
    np.testing.assert_array_equal(ix.dbscan_predict(lab, nc, eps, q), O.dbscan_predict(fit, olab, onc, eps, q))


def test_many_clusters_and_borders_equal_oracle():
    """20,000 rows: four touching Gaussian blobs and uniform noise, eps 0.45,
    min_samples 8 -- 37 clusters, most of them small groups on the blobs'
    fringes, and many non-core rows next to more than one cluster.  The
    device union-find over the core rows' edges, the numbering by smallest
    core row and the lowest-cluster rule for the other rows must give the
    reference loop's labels exactly."""
    rng = np.random.default_rng(63)
    parts = [rng.normal(c, 1.0, size=(4000, 4))
             for c in ([0, 0, 0, 0], [3.2, 0, 0, 0], [0, 6, 0, 0], [9, 9, 0, 0])]
    parts.append(rng.uniform(-6, 14, size=(4000, 4)))
    x = np.concatenate(parts)
    x = x[rng.permutation(len(x))]
    lab, nc = S.Index(x).dbscan_fit(0.45, 8)
    olab, onc = O.dbscan_fit(x, 0.45, 8)
    assert nc == onc and nc > 20
    np.testing.assert_array_equal(lab, olab)


@pytest.mark.parametrize("metric,p,eps", [(S.SC_EUCLIDIAN, 2, 2.2), (S.SC_MANHATTAN, 1, 9.0),
                                          (S.SC_MINKOWSKI, 3, 1.7)])
def test_blobs_equal_oracle(metric, p, eps):
    xtr, _, xq, _ = synth.knn_mixture(seed=7, n=3000, q=300, d=8)
    ix = S.Index(xtr)
    lab, nc = ix.dbscan_fit(eps, 6, metric, p)
    olab, onc = O.dbscan_fit(xtr, eps, 6, metric, p)
    assert nc == onc and nc >= 1
    np.testing.assert_array_equal(lab, olab)
    np.testing.assert_array_equal(ix.dbscan_predict(lab, nc, eps, xq, metric, p),
                                  O.dbscan_predict(xtr, olab, onc, eps, xq, metric, p))


def test_f32_mirror_matches_capi():
    x = np.load(os.path.join(G, "digits.npz"))["x"][:800].astype(np.float32)
    m = sc.DBSCAN.fit(x, sc.DBSCANParameters().with_eps(20.0).with_min_samples(8))
    lab, nc = S.Index(x).dbscan_fit(20.0, 8)
    np.testing.assert_array_equal(m.cluster_labels, lab)
    assert m.num_classes == nc
