"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker and the single-core CPU
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module; the product package
``smartcore_amd`` never does.

The C restatement lives in ``oracle/oracle.c`` and follows smartcore v0.4.0
(``src/metrics/distance``, ``src/algorithm/sort/heap_select.rs``,
``src/algorithm/neighbour/{linear_search,bbd_tree}.rs``, ``src/neighbors``,
``src/cluster/kmeans.rs``) line by line.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F64, F32 = 0, 1
EUCLIDIAN, MANHATTAN, MINKOWSKI = 0, 1, 2

_lib = None


def build() -> None:
    """Compile liboracle.so from oracle.c (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, sz, dbl, i32, u32 = C.c_void_p, C.c_size_t, C.c_double, C.c_int, C.c_uint
        L.or_squared_distance_f64.restype = dbl
        L.or_squared_distance_f64.argtypes = [vp, vp, sz]
        L.or_squared_distance_f32.restype = dbl
        L.or_squared_distance_f32.argtypes = [vp, vp, sz]
        L.or_distance.restype = dbl
        L.or_distance.argtypes = [i32, u32, i32, vp, vp, sz, C.POINTER(C.c_int)]
        L.or_heap_init.argtypes = [vp, vp, sz]
        L.or_heap_add.argtypes = [vp, dbl, C.c_int64]
        L.or_heap_heapify.argtypes = [vp]
        L.or_linear_find.restype = i32
        L.or_linear_find.argtypes = [vp, i32, sz, sz, vp, sz, i32, u32, vp, vp, C.POINTER(C.c_size_t)]
        L.or_linear_find_via_add.restype = i32
        L.or_linear_find_via_add.argtypes = L.or_linear_find.argtypes
        L.or_linear_find_batch.restype = i32
        L.or_linear_find_batch.argtypes = [vp, i32, sz, sz, vp, sz, sz, i32, u32, vp, vp, vp]
        L.or_linear_find_radius.restype = i32
        L.or_linear_find_radius.argtypes = [vp, i32, sz, sz, vp, dbl, i32, u32, vp, vp, sz,
                                            C.POINTER(C.c_size_t)]
        L.or_calc_weights.argtypes = [i32, vp, sz, vp]
        L.or_knn_classifier_predict_row.restype = sz
        L.or_knn_classifier_predict_row.argtypes = [vp, vp, sz, vp, sz, i32]
        L.or_knn_regressor_predict_row.restype = dbl
        L.or_knn_regressor_predict_row.argtypes = [vp, vp, sz, vp, i32]
        L.or_bbd_new.restype = vp
        L.or_bbd_new.argtypes = [vp, i32, sz, sz]
        L.or_bbd_free.argtypes = [vp]
        L.or_bbd_num_nodes.restype = sz
        L.or_bbd_num_nodes.argtypes = [vp]
        L.or_bbd_clustering.restype = dbl
        L.or_bbd_clustering.argtypes = [vp, vp, sz, vp, vp, vp]
        L.or_kmeans_means_from_labels.argtypes = [vp, i32, sz, sz, vp, sz, vp, vp]
        L.or_kmeans_fit.restype = i32
        L.or_kmeans_fit.argtypes = [vp, i32, sz, sz, sz, sz, i32, vp, vp, vp,
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
        L.or_lloyd_step_brute.restype = dbl
        L.or_lloyd_step_brute.argtypes = [vp, i32, sz, sz, vp, sz, vp, vp, vp]
        L.or_kmeans_predict.argtypes = [vp, i32, sz, sz, vp, sz, vp]
        u64 = C.c_uint64
        L.or_pcg32_output.restype = C.c_uint32
        L.or_pcg32_output.argtypes = [u64]
        L.or_smallrng_seed.argtypes = [u64, vp]
        L.or_smallrng_next.restype = u64
        L.or_smallrng_next.argtypes = [vp]
        L.or_smallrng_gen_range.restype = u64
        L.or_smallrng_gen_range.argtypes = [vp, u64]
        L.or_smallrng_gen_f64.restype = dbl
        L.or_smallrng_gen_f64.argtypes = [vp]
        L.or_kmeans_plus_plus.argtypes = [vp, i32, sz, sz, sz, i32, u64, vp]
        L.or_dbscan_fit.restype = i32
        L.or_dbscan_fit.argtypes = [vp, i32, sz, sz, dbl, sz, i32, u32, vp,
                                    C.POINTER(C.c_size_t)]
        L.or_dbscan_predict.restype = i32
        L.or_dbscan_predict.argtypes = [vp, i32, sz, sz, vp, sz, dbl, i32, u32, vp, sz, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dt(a: np.ndarray) -> int:
    if a.dtype == np.float64:
        return F64
    if a.dtype == np.float32:
        return F32
    raise TypeError(f"oracle supports float64/float32, got {a.dtype}")


class OracleError(Exception):
    def __init__(self, code: int):
        super().__init__(f"oracle status {code}")
        self.code = code


def distance(metric: int, a, b, p: int = 2) -> float:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    err = C.c_int(0)
    v = lib().or_distance(metric, p, _dt(a), _p(a), _p(b), a.shape[0], C.byref(err))
    if err.value:
        raise OracleError(err.value)
    return v


def squared_distance(a, b) -> float:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    f = lib().or_squared_distance_f64 if a.dtype == np.float64 else lib().or_squared_distance_f32
    return f(_p(a), _p(b), a.shape[0])


class _Point(C.Structure):
    _fields_ = [("distance", C.c_double), ("index", C.c_int64)]


class _Heap(C.Structure):
    _fields_ = [("k", C.c_size_t), ("n", C.c_size_t), ("sorted", C.c_int),
                ("len", C.c_size_t), ("heap", C.c_void_p)]


class HeapSelection:
    """HeapSelection<f64> of heap_select.rs driven through the oracle."""

    def __init__(self, k: int):
        self._store = (_Point * max(k, 1))()
        self._h = _Heap()
        lib().or_heap_init(C.byref(self._h), C.cast(self._store, C.c_void_p), k)

    def add(self, v: float, index: int = -1) -> None:
        lib().or_heap_add(C.byref(self._h), float(v), index)

    def heapify(self) -> None:
        lib().or_heap_heapify(C.byref(self._h))

    @property
    def n(self) -> int:
        return self._h.n

    def get(self):
        return [self._store[i].distance for i in range(self._h.len)]


def linear_find(data: np.ndarray, query: np.ndarray, k: int, metric: int = EUCLIDIAN,
                p: int = 2, via_add: bool = False):
    """LinearKNNSearch::find: (indices, distances) in heap-array order."""
    data = np.ascontiguousarray(data)
    query = np.ascontiguousarray(query, dtype=data.dtype)
    n, d = data.shape
    idx = np.full(max(k, 1), -1, np.int64)
    dist = np.full(max(k, 1), np.nan)
    cnt = C.c_size_t(0)
    f = lib().or_linear_find_via_add if via_add else lib().or_linear_find
    rc = f(_p(data), _dt(data), n, d, _p(query), k, metric, p, _p(idx), _p(dist), C.byref(cnt))
    if rc:
        raise OracleError(rc)
    return idx[: cnt.value].copy(), dist[: cnt.value].copy()


def linear_find_batch(data: np.ndarray, queries: np.ndarray, k: int, metric: int = EUCLIDIAN,
                      p: int = 2):
    data = np.ascontiguousarray(data)
    queries = np.ascontiguousarray(queries, dtype=data.dtype)
    n, d = data.shape
    q = queries.shape[0]
    idx = np.empty((q, k), np.int64)
    dist = np.empty((q, k), np.float64)
    cnt = np.empty(q, np.uint32)
    rc = lib().or_linear_find_batch(_p(data), _dt(data), n, d, _p(queries), q, k, metric, p,
                                    _p(idx), _p(dist), _p(cnt))
    if rc:
        raise OracleError(rc)
    return idx, dist, cnt


def linear_find_radius(data: np.ndarray, query: np.ndarray, radius: float,
                       metric: int = EUCLIDIAN, p: int = 2):
    data = np.ascontiguousarray(data)
    query = np.ascontiguousarray(query, dtype=data.dtype)
    n, d = data.shape
    idx = np.empty(n, np.int64)
    dist = np.empty(n, np.float64)
    cnt = C.c_size_t(0)
    rc = lib().or_linear_find_radius(_p(data), _dt(data), n, d, _p(query), float(radius), metric,
                                     p, _p(idx), _p(dist), n, C.byref(cnt))
    if rc:
        raise OracleError(rc)
    return idx[: cnt.value].copy(), dist[: cnt.value].copy()


def calc_weights(weight: int, dist) -> np.ndarray:
    dist = np.ascontiguousarray(dist, dtype=np.float64)
    w = np.empty_like(dist)
    lib().or_calc_weights(weight, _p(dist), dist.shape[0], _p(w))
    return w


def classifier_predict_row(idx, dist, y_class, n_classes: int, weight: int) -> int:
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    dist = np.ascontiguousarray(dist, dtype=np.float64)
    y_class = np.ascontiguousarray(y_class, dtype=np.int64)
    return lib().or_knn_classifier_predict_row(_p(idx), _p(dist), idx.shape[0], _p(y_class),
                                               n_classes, weight)


def regressor_predict_row(idx, dist, y, weight: int) -> float:
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    dist = np.ascontiguousarray(dist, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().or_knn_regressor_predict_row(_p(idx), _p(dist), idx.shape[0], _p(y), weight)


class BBDTree:
    """BBDTree of bbd_tree.rs (built once, clustering per Lloyd step)."""

    def __init__(self, data: np.ndarray):
        self.data = np.ascontiguousarray(data)
        n, d = self.data.shape
        self.n, self.d = n, d
        self._t = lib().or_bbd_new(_p(self.data), _dt(self.data), n, d)

    def __del__(self):
        if getattr(self, "_t", None):
            lib().or_bbd_free(self._t)
            self._t = None

    def num_nodes(self) -> int:
        return lib().or_bbd_num_nodes(self._t)

    def clustering(self, centroids: np.ndarray, membership: np.ndarray | None = None):
        centroids = np.ascontiguousarray(centroids, dtype=np.float64)
        k = centroids.shape[0]
        sums = np.zeros((k, self.d))
        counts = np.zeros(k, np.uint64)
        if membership is None:
            membership = np.zeros(self.n, np.uint32)
        dist = lib().or_bbd_clustering(self._t, _p(centroids), k, _p(sums), _p(counts),
                                       _p(membership))
        return dist, sums, counts, membership


def lloyd_step_brute(data: np.ndarray, centroids: np.ndarray):
    data = np.ascontiguousarray(data)
    centroids = np.ascontiguousarray(centroids, dtype=np.float64)
    n, d = data.shape
    k = centroids.shape[0]
    sums = np.zeros((k, d))
    counts = np.zeros(k, np.uint64)
    memb = np.zeros(n, np.uint32)
    dist = lib().or_lloyd_step_brute(_p(data), _dt(data), n, d, _p(centroids), k, _p(sums),
                                     _p(counts), _p(memb))
    return dist, sums, counts, memb


def means_from_labels(data: np.ndarray, y: np.ndarray, k: int):
    data = np.ascontiguousarray(data)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    n, d = data.shape
    c = np.zeros((k, d))
    size = np.zeros(k, np.uint64)
    lib().or_kmeans_means_from_labels(_p(data), _dt(data), n, d, _p(y), k, _p(c), _p(size))
    return c, size


def kmeans_fit(data: np.ndarray, init_centroids: np.ndarray, max_iter: int, use_bbd: bool = True):
    """KMeans::fit's loop (kmeans.rs:294-310) from given initial centroids."""
    data = np.ascontiguousarray(data)
    n, d = data.shape
    cent = np.array(init_centroids, dtype=np.float64, order="C", copy=True)
    k = cent.shape[0]
    labels = np.zeros(n, np.uint32)
    size = np.zeros(k, np.uint64)
    dist = C.c_double(0)
    iters = C.c_uint32(0)
    rc = lib().or_kmeans_fit(_p(data), _dt(data), n, d, k, max_iter, int(use_bbd), _p(cent),
                             _p(labels), _p(size), C.byref(dist), C.byref(iters))
    if rc:
        raise OracleError(rc)
    return cent, labels, size, dist.value, iters.value


class SmallRng:
    """rand 0.8.5's SmallRng (xoshiro256++) as SmallRng::seed_from_u64 seeds
    it (rand_core 0.6's PCG32 seed fill); `state` sets the four words
    directly, as xoshiro256++'s from_seed does."""

    def __init__(self, seed: int | None = None, state=None):
        self.s = np.zeros(4, np.uint64)
        if state is not None:
            self.s[:] = np.asarray(state, dtype=np.uint64)
        else:
            lib().or_smallrng_seed(int(seed or 0), _p(self.s))

    def next_u64(self) -> int:
        return int(lib().or_smallrng_next(_p(self.s)))

    def gen_range(self, n: int) -> int:
        return int(lib().or_smallrng_gen_range(_p(self.s), n))

    def gen_f64(self) -> float:
        return float(lib().or_smallrng_gen_f64(_p(self.s)))


def pcg32_output(state: int) -> int:
    return int(lib().or_pcg32_output(state))


def kmeans_plus_plus(data: np.ndarray, k: int, seed: int | None = None) -> np.ndarray:
    """KMeans::kmeans_plus_plus (kmeans.rs:354-413): initial labels."""
    data = np.ascontiguousarray(data)
    n, d = data.shape
    y = np.zeros(max(n, 1), np.uint32)
    lib().or_kmeans_plus_plus(_p(data), _dt(data), n, d, k, int(seed is not None), int(seed or 0),
                              _p(y))
    return y[:n]


def kmeans_predict(data: np.ndarray, centroids: np.ndarray) -> np.ndarray:
    data = np.ascontiguousarray(data)
    centroids = np.ascontiguousarray(centroids, dtype=np.float64)
    n, d = data.shape
    labels = np.zeros(n, np.uint32)
    lib().or_kmeans_predict(_p(data), _dt(data), n, d, _p(centroids), centroids.shape[0],
                            _p(labels))
    return labels


def dbscan_fit(data: np.ndarray, eps: float, min_samples: int, metric: int = EUCLIDIAN,
               p: int = 2):
    """DBSCAN::fit (dbscan.rs:274-354): (cluster_labels int32, num_classes)."""
    data = np.ascontiguousarray(data)
    n, d = data.shape
    labels = np.zeros(max(n, 1), np.int32)
    nc = C.c_size_t(0)
    rc = lib().or_dbscan_fit(_p(data), _dt(data), n, d, eps, min_samples, metric, p,
                             _p(labels), C.byref(nc))
    if rc:
        raise OracleError(rc)
    return labels[:n], nc.value


def dbscan_predict(data: np.ndarray, labels: np.ndarray, num_classes: int, eps: float,
                   queries: np.ndarray, metric: int = EUCLIDIAN, p: int = 2) -> np.ndarray:
    """DBSCAN::predict (dbscan.rs:358-397)."""
    data = np.ascontiguousarray(data)
    queries = np.ascontiguousarray(queries, dtype=data.dtype)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    n, d = data.shape
    out = np.zeros(max(queries.shape[0], 1), np.int32)
    rc = lib().or_dbscan_predict(_p(data), _dt(data), n, d, _p(labels), num_classes, eps,
                                 metric, p, _p(queries), queries.shape[0], _p(out))
    if rc:
        raise OracleError(rc)
    return out[:queries.shape[0]]
