"""Seeded synthetic workloads for the kNN and k-means paths.

Host generators use numpy's PCG64 (``np.random.default_rng(seed)``); the
large device-side generators use torch's Philox generator on the GPU.  Every
array is produced in its stored dtype, so f32 rows are exactly the values the
f64 forms widen.

Workloads (sizes are parameters; the defaults are the named workloads):
  knn_mixture      KNNClassifier data: 10 class centres ~ N(0, 4^2 I), points =
                   centre + N(0, I), class drawn uniformly, labels i32; the
                   queries come from the same mixture (10,000 x 32 train,
                   2,000 queries, f64).
  kmeans_blobs     64 centres ~ U(-10, 10)^d, per-blob std ~ U(0.5, 2), blob
                   sizes multinomial with Dirichlet(1) weights, rows shuffled;
                   initial centroids are k distinct rows chosen by the seed
                   (1,000,000 x 64, k = 64, f32).
  lognormal_rows   rows ~ N(0, I) scaled by per-row norms ~ LogNormal(0, 0.5)
                   (4,000,000 x 128 reference, 65,536 queries, f32).
"""
from __future__ import annotations

import numpy as np


def knn_mixture(seed: int = 1, n: int = 10_000, q: int = 2_000, d: int = 32, classes: int = 10,
                dtype=np.float64):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 4.0, size=(classes, d))
    ytr = rng.integers(0, classes, size=n).astype(np.int32)
    xtr = centres[ytr] + rng.normal(size=(n, d))
    yq = rng.integers(0, classes, size=q).astype(np.int32)
    xq = centres[yq] + rng.normal(size=(q, d))
    return xtr.astype(dtype), ytr, xq.astype(dtype), yq


def kmeans_blobs(seed: int = 2, n: int = 1_000_000, d: int = 64, blobs: int = 64, k: int = 64,
                 dtype=np.float32):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-10.0, 10.0, size=(blobs, d))
    std = rng.uniform(0.5, 2.0, size=blobs)
    w = rng.dirichlet(np.ones(blobs))
    sizes = rng.multinomial(n, w)
    lab = np.repeat(np.arange(blobs), sizes)
    rng.shuffle(lab)
    x = np.empty((n, d), dtype=dtype)
    step = 1 << 18
    for s in range(0, n, step):
        e = min(n, s + step)
        lb = lab[s:e]
        x[s:e] = (centres[lb] + std[lb, None] * rng.normal(size=(e - s, d))).astype(dtype)
    init = rng.choice(n, size=k, replace=False)
    init_centroids = x[np.sort(init)].astype(np.float64)
    return x, lab.astype(np.int32), init_centroids


def lognormal_rows(seed: int, n: int, d: int, dtype=np.float32):
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(n, d))
    r = rng.lognormal(0.0, 0.5, size=(n, 1))
    return (z * r).astype(dtype)


def lognormal_rows_device(seed: int, n: int, d: int, device="cuda"):
    """Same distribution generated on the GPU (torch Philox), f32."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    x = torch.empty((n, d), dtype=torch.float32, device=device)
    step = 1 << 20
    for s in range(0, n, step):
        e = min(n, s + step)
        z = torch.randn((e - s, d), generator=g, device=device, dtype=torch.float32)
        r = torch.exp(0.5 * torch.randn((e - s, 1), generator=g, device=device, dtype=torch.float32))
        x[s:e] = z * r
    return x


def kmeans_batch_problems(seed: int = 4, count: int = 512, n_lo: int = 1_000, n_hi: int = 200_000,
                          dims=(8, 16, 32, 64), k_lo: int = 4, k_hi: int = 128, dtype=np.float32):
    """Uneven k-means problems: n ~ log-uniform[n_lo, n_hi], d from `dims`,
    k ~ U{k_lo..k_hi}; rows are Gaussian blobs as in knn_mixture (k centres
    ~ N(0, 4^2 I), unit noise, centre drawn uniformly per row); initial
    centroids are k distinct rows chosen by the seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        d = int(rng.choice(dims))
        k = int(rng.integers(k_lo, k_hi + 1))
        k = min(k, n)
        centres = rng.normal(0.0, 4.0, size=(k, d))
        lab = rng.integers(0, k, size=n)
        x = (centres[lab] + rng.normal(size=(n, d))).astype(dtype)
        init = x[np.sort(rng.choice(n, size=k, replace=False))].astype(np.float64)
        out.append((x, init))
    return out
