"""Writes the golden fixtures of tests/test_golden.py.

Run from the repository root in a container that holds the reference
(`python tests/golden/make_golden.py`).  It reads two data files the
reference ships with its datasets -- `src/dataset/digits.xy` and
`src/dataset/iris.xy`, whose layout is dataset/mod.rs:86-118: u64
num_features, u64 num_samples, then the samples' f32 values row by row and
the f32 targets -- and stores them compactly.  Expected outputs come from
the oracle (the CPU restatement of the reference) on those data and on
seeded synthetic inputs:

  digits.npz    the 1797 x 64 digits (integer pixels 0..16) and labels
  iris.npz      the 150 x 4 iris measurements and labels
  expected.npz  heap-order neighbours on digits (Euclidean k=5, Manhattan
                k=7), a k-means fit on digits from its first 10 rows, the
                first 256 queries of the 10,000 x 32 classifier workload
                (neighbours, distance bits, predicted labels), and a
                20,000 x 16, k=8 k-means fit (labels, sizes, iterations,
                centroids)
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/src/dataset"


def read_xy(path):
    b = open(path, "rb").read()
    nf, ns = struct.unpack_from("<QQ", b, 0)
    x = np.frombuffer(b, dtype="<f4", count=ns * nf, offset=16).reshape(ns, nf)
    y = np.frombuffer(b, dtype="<f4", count=ns, offset=16 + 4 * ns * nf)
    assert 16 + 4 * ns * (nf + 1) == len(b)
    return x.copy(), y.copy()


def main():
    import oracle as O
    from smartcore_amd import synth

    dx, dy = read_xy(os.path.join(REF, "digits.xy"))
    assert dx.shape == (1797, 64) and np.all(dx == np.round(dx)) and dx.min() >= 0 and dx.max() <= 16
    np.savez_compressed(os.path.join(HERE, "digits.npz"), x=dx.astype(np.uint8),
                        y=dy.astype(np.uint8))
    ix, iy = read_xy(os.path.join(REF, "iris.xy"))
    np.savez_compressed(os.path.join(HERE, "iris.npz"), x=ix, y=iy)

    out = {}
    d32 = dx.astype(np.float32)
    i, dd, c = O.linear_find_batch(d32, d32[:256], 5, O.EUCLIDIAN)
    out.update(digits_knn_idx=i, digits_knn_dist=dd.view(np.uint64), digits_knn_count=c)
    i, dd, c = O.linear_find_batch(d32, d32[:64], 7, O.MANHATTAN)
    out.update(digits_l1_idx=i, digits_l1_dist=dd.view(np.uint64), digits_l1_count=c)
    cent, lab, size, dist, it = O.kmeans_fit(d32, d32[:10].astype(np.float64), 30, use_bbd=True)
    out.update(digits_km_labels=lab, digits_km_size=size, digits_km_iters=np.array(it),
               digits_km_centroids=cent)

    xtr, ytr, xq, yq = synth.knn_mixture(seed=1)
    q = xq[:256]
    i, dd, c = O.linear_find_batch(xtr, q, 5, O.EUCLIDIAN)
    pred = np.array([O.classifier_predict_row(i[r, :c[r]], dd[r, :c[r]], ytr, 10, 0)
                     for r in range(len(q))], dtype=np.int32)
    out.update(c1_idx=i, c1_dist=dd.view(np.uint64), c1_pred=pred)

    x, _, init = synth.kmeans_blobs(seed=5, n=20_000, d=16, blobs=8, k=8)
    cent, lab, size, dist, it = O.kmeans_fit(x, init, 20, use_bbd=True)
    out.update(km_labels=lab, km_size=size, km_iters=np.array(it), km_centroids=cent,
               km_init=init)
    np.savez_compressed(os.path.join(HERE, "expected.npz"), **out)
    for k, v in out.items():
        print(k, v.shape, v.dtype)


if __name__ == "__main__":
    main()
