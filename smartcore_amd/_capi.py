"""ctypes binding of libsmartcore_amd.so, the C ABI in include/smartcore_amd.h.

The library is loaded from this package's directory (built in-tree by
``smartcore_amd.build``).  There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmartcore_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "smartcore_amd.h")

SC_OK = 0
SC_FIT_FAILED = 1
SC_PREDICT_FAILED = 2
SC_TRANSFORM_FAILED = 3
SC_FIND_FAILED = 4
SC_DECOMPOSITION_FAILED = 5
SC_SOLUTION_FAILED = 6
SC_PARAMETERS_ERROR = 7
SC_INVALID_STATE = 8
SC_ERR_PANIC = 64
SC_ERR_DEVICE = 65
SC_ERR_OUT_OF_MEMORY = 66
SC_ERR_UNSUPPORTED = 67

SC_F64, SC_F32 = 0, 1
SC_EUCLIDIAN, SC_MANHATTAN, SC_MINKOWSKI = 0, 1, 2
SC_PLAIN_FP64, SC_MIXED = 0, 1

# smartcore's FailedError Display strings (src/error/mod.rs:108-121)
_KIND_TEXT = {
    SC_FIT_FAILED: "Fit failed",
    SC_PREDICT_FAILED: "Predict failed",
    SC_TRANSFORM_FAILED: "Transform failed",
    SC_FIND_FAILED: "Find failed",
    SC_DECOMPOSITION_FAILED: "Decomposition failed",
    SC_SOLUTION_FAILED: "Can't find solution",
    SC_PARAMETERS_ERROR: "Error in input, check parameters",
    SC_INVALID_STATE: "Invalid state, this should never happen",
}


class Failed(Exception):
    """smartcore's Failed{err, msg}; str() is "{err}: {msg}" (error/mod.rs:124-127)."""

    def __init__(self, code: int, msg: str):
        self.code = code
        self.msg = msg
        kind = _KIND_TEXT.get(code)
        super().__init__(f"{kind}: {msg}" if kind else msg)


class Panic(Failed):
    """A condition the reference turns into a panic (dimension mismatch, p < 1)."""


class DeviceError(Failed):
    """HIP runtime failure or device memory exhaustion."""


class sc_stats(C.Structure):
    _fields_ = [
        ("queries", C.c_uint64),
        ("candidates", C.c_uint64),
        ("overflow_queries", C.c_uint64),
        ("shortlist_rescored", C.c_uint64),
        ("uncertified", C.c_uint64),
        ("launch_waves", C.c_uint64),
        ("launch_groups_per_cu", C.c_uint64),
        ("upload_groups", C.c_uint64),
    ]


_lib = None


def declared_symbols() -> list[str]:
    """Entry points declared in include/smartcore_amd.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m smartcore_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, sz, i32, u16, dbl = C.c_void_p, C.c_size_t, C.c_int, C.c_uint16, C.c_double
    L.sc_index_create.argtypes = [vp, i32, sz, sz, C.POINTER(vp)]
    L.sc_index_create_device.argtypes = [vp, i32, sz, sz, C.POINTER(vp)]
    L.sc_index_destroy.argtypes = [vp]
    L.sc_index_rows.argtypes = [vp]
    L.sc_index_rows.restype = sz
    L.sc_index_dim.argtypes = [vp]
    L.sc_index_dim.restype = sz
    L.sc_knn_search.argtypes = [vp, vp, sz, sz, i32, u16, i32, vp, vp, vp]
    L.sc_knn_search_device.argtypes = [vp, vp, sz, sz, i32, u16, i32, vp, vp, vp, vp]
    L.sc_radius_search.argtypes = [vp, vp, sz, dbl, i32, u16, C.POINTER(vp)]
    L.sc_radius_result_total.argtypes = [vp]
    L.sc_radius_result_total.restype = sz
    L.sc_radius_result_copy.argtypes = [vp, vp, vp, vp]
    L.sc_radius_result_free.argtypes = [vp]
    L.sc_radius_result_free.restype = None
    L.sc_kmeans_step.argtypes = [vp, vp, sz, vp, vp, vp, C.POINTER(dbl)]
    L.sc_kmeans_run.argtypes = [vp, sz, sz, i32, vp, vp, vp, C.POINTER(dbl), C.POINTER(C.c_uint32)]
    L.sc_kmeans_predict.argtypes = [vp, i32, sz, sz, vp, sz, vp]
    L.sc_dbscan_fit.argtypes = [vp, dbl, sz, i32, u16, vp, C.POINTER(sz)]
    L.sc_dbscan_predict.argtypes = [vp, vp, sz, dbl, i32, u16, vp, sz, vp]
    L.sc_last_error.argtypes = [C.c_char_p, sz]
    L.sc_last_error.restype = sz
    L.sc_last_stats.argtypes = [C.POINTER(sc_stats)]
    for name in declared_symbols():
        fn = getattr(L, name)
        if fn.restype is C.c_int and name not in ("sc_radius_result_free",):
            fn.restype = C.c_int
    _lib = L
    return L


def last_error() -> str:
    n = lib().sc_last_error(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().sc_last_error(buf, n + 1)
    return buf.value.decode()


def last_stats() -> dict:
    st = sc_stats()
    lib().sc_last_stats(C.byref(st))
    return {f: getattr(st, f) for f, _ in sc_stats._fields_}


def check(rc: int) -> None:
    if rc == SC_OK:
        return
    msg = last_error()
    if rc == SC_ERR_PANIC:
        raise Panic(rc, msg)
    if rc in (SC_ERR_DEVICE, SC_ERR_OUT_OF_MEMORY):
        raise DeviceError(rc, msg)
    raise Failed(rc, msg)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float64:
        return SC_F64
    if dt == np.float32:
        return SC_F32
    raise TypeError(f"rows must be float64 or float32, got {dt}")


class sc_knn_block(C.Structure):
    _fields_ = [("queries", C.c_void_p), ("q", C.c_size_t), ("k", C.c_size_t),
                ("idx", C.c_void_p), ("dist", C.c_void_p), ("count", C.c_void_p)]


class Index:
    """Device-resident reference rows (an sc_index handle)."""

    def __init__(self, rows: np.ndarray | None = None, *, device_ptr: int | None = None,
                 n: int = 0, d: int = 0, dtype=np.float64):
        h = C.c_void_p()
        if device_ptr is not None:
            self.dtype = np.dtype(dtype)
            check(lib().sc_index_create_device(C.c_void_p(device_ptr), dtype_code(self.dtype), n, d,
                                               C.byref(h)))
        else:
            rows = np.ascontiguousarray(rows)
            if rows.ndim != 2:
                raise ValueError("rows must be 2-D")
            self.dtype = rows.dtype
            n, d = rows.shape
            check(lib().sc_index_create(_ptr(rows), dtype_code(rows.dtype), n, d, C.byref(h)))
        self._h = h
        self.n, self.d = n, d

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().sc_index_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def knn(self, queries: np.ndarray, k: int, metric: int = SC_EUCLIDIAN, p: int = 2,
            mode: int = SC_PLAIN_FP64):
        """sc_knn_search: (idx[q,k], dist[q,k], count[q]) in heap-array order."""
        queries = np.ascontiguousarray(queries, dtype=self.dtype)
        if queries.ndim == 1:
            queries = queries[None, :]
        q = queries.shape[0]
        if q and queries.shape[1] != self.d:
            raise Panic(SC_ERR_PANIC, "Input vector sizes are different.")
        kk = max(int(k), 0)
        idx = np.empty((q, max(kk, 1)), np.int64)
        dist = np.empty((q, max(kk, 1)), np.float64)
        cnt = np.empty(max(q, 1), np.uint32)
        check(lib().sc_knn_search(self._h, _ptr(queries), q, kk, metric, p, mode, _ptr(idx),
                                  _ptr(dist), _ptr(cnt)))
        return idx[:, :kk], dist[:, :kk], cnt[:q]

    def knn_batch(self, blocks, metric: int = SC_EUCLIDIAN, p: int = 2,
                  mode: int = SC_PLAIN_FP64):
        """sc_knn_search_batch over [(queries, k), ...]: a list of
        (idx, dist, count) per block."""
        arr = (sc_knn_block * max(len(blocks), 1))()
        keep, outs = [], []
        for i, (qs, k) in enumerate(blocks):
            qs = np.ascontiguousarray(qs, dtype=self.dtype)
            q = qs.shape[0]
            idx = np.empty((q, max(k, 1)), np.int64)
            dist = np.empty((q, max(k, 1)), np.float64)
            cnt = np.empty(max(q, 1), np.uint32)
            keep += [qs, idx, dist, cnt]
            outs.append((idx, dist, cnt, q, k))
            arr[i] = sc_knn_block(qs.ctypes.data, q, k, idx.ctypes.data, dist.ctypes.data,
                                  cnt.ctypes.data)
        L = lib()
        L.sc_knn_search_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                          C.c_uint16, C.c_int]
        check(L.sc_knn_search_batch(self._h, C.cast(arr, C.c_void_p), len(blocks), metric, p, mode))
        return [(i[:, :k], d[:, :k], c[:q]) for i, d, c, q, k in outs]

    def knn_device(self, q_ptr: int, q: int, k: int, idx_ptr: int, dist_ptr: int, cnt_ptr: int,
                   metric: int = SC_EUCLIDIAN, p: int = 2, mode: int = SC_PLAIN_FP64,
                   stream: int = 0) -> None:
        check(lib().sc_knn_search_device(self._h, C.c_void_p(q_ptr), q, k, metric, p, mode,
                                         C.c_void_p(idx_ptr), C.c_void_p(dist_ptr),
                                         C.c_void_p(cnt_ptr), C.c_void_p(stream)))

    def radius(self, queries: np.ndarray, radius: float, metric: int = SC_EUCLIDIAN, p: int = 2):
        """sc_radius_search: list of (idx, dist) arrays, one per query, index order."""
        queries = np.ascontiguousarray(queries, dtype=self.dtype)
        if queries.ndim == 1:
            queries = queries[None, :]
        q = queries.shape[0]
        r = C.c_void_p()
        check(lib().sc_radius_search(self._h, _ptr(queries), q, float(radius), metric, p,
                                     C.byref(r)))
        try:
            total = lib().sc_radius_result_total(r)
            off = np.empty(q + 1, np.uint64)
            idx = np.empty(max(total, 1), np.int64)
            dist = np.empty(max(total, 1), np.float64)
            check(lib().sc_radius_result_copy(r, _ptr(off), _ptr(idx), _ptr(dist)))
        finally:
            lib().sc_radius_result_free(r)
        return [(idx[off[i]:off[i + 1]].copy(), dist[off[i]:off[i + 1]].copy()) for i in range(q)]

    def dbscan_fit(self, eps: float, min_samples: int, metric: int = SC_EUCLIDIAN, p: int = 2):
        """sc_dbscan_fit: (cluster_labels int32[n], num_classes)."""
        labels = np.empty(max(self.n, 1), np.int32)
        nc = C.c_size_t()
        check(lib().sc_dbscan_fit(self._h, float(eps), min_samples, metric, p, _ptr(labels),
                                  C.byref(nc)))
        return labels[: self.n], nc.value

    def dbscan_predict(self, labels: np.ndarray, num_classes: int, eps: float,
                       queries: np.ndarray, metric: int = SC_EUCLIDIAN, p: int = 2):
        """sc_dbscan_predict: int32[q] (cluster + 1, or 0)."""
        queries = np.ascontiguousarray(queries, dtype=self.dtype)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        q = queries.shape[0]
        out = np.empty(max(q, 1), np.int32)
        check(lib().sc_dbscan_predict(self._h, _ptr(labels), num_classes, float(eps), metric, p,
                                      _ptr(queries), q, _ptr(out)))
        return out[:q]

    def kmeans_step(self, centroids: np.ndarray):
        """sc_kmeans_step: (distortion, sums[k,d], counts[k], membership[n])."""
        centroids = np.ascontiguousarray(centroids, dtype=np.float64)
        k = centroids.shape[0]
        sums = np.zeros((k, self.d))
        counts = np.zeros(k, np.uint64)
        memb = np.zeros(max(self.n, 1), np.uint32)
        dist = C.c_double()
        check(lib().sc_kmeans_step(self._h, _ptr(centroids), k, _ptr(sums), _ptr(counts),
                                   _ptr(memb), C.byref(dist)))
        return dist.value, sums, counts, memb[: self.n]

    def kmeans_plus_plus(self, k: int, seed: int | None = None) -> np.ndarray:
        """sc_kmeans_plus_plus: initial labels (KMeans::kmeans_plus_plus)."""
        L = lib()
        L.sc_kmeans_plus_plus.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_uint64,
                                          C.c_void_p]
        y = np.zeros(max(self.n, 1), np.uint32)
        check(L.sc_kmeans_plus_plus(self._h, k, int(seed is not None), int(seed or 0), _ptr(y)))
        return y[: self.n]

    def kmeans_run(self, init_centroids: np.ndarray, max_iter: int, mode: int = SC_PLAIN_FP64):
        """sc_kmeans_run: (centroids, labels, size, distortion, iters)."""
        cent = np.array(init_centroids, dtype=np.float64, order="C", copy=True)
        k = cent.shape[0] if cent.ndim == 2 else 0
        if cent.ndim != 2:
            cent = np.zeros((max(k, 1), self.d))
        labels = np.zeros(max(self.n, 1), np.uint32)
        size = np.zeros(max(k, 1), np.uint64)
        dist = C.c_double()
        iters = C.c_uint32()
        check(lib().sc_kmeans_run(self._h, k, max_iter, mode, _ptr(cent), _ptr(labels), _ptr(size),
                                  C.byref(dist), C.byref(iters)))
        return cent, labels[: self.n], size[:k], dist.value, iters.value


class sc_server_config(C.Structure):
    _fields_ = [("max_batch_queries", C.c_size_t), ("max_inflight", C.c_size_t),
                ("hbm_budget_bytes", C.c_size_t), ("batch_window_us", C.c_uint32)]


class sc_server_info(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("completed", "batches", "refused", "queries", "pending",
                                          "hbm_in_use", "hbm_peak", "hbm_budget")]


class Server:
    """sc_server_*: continuous-batching kNN server over one resident index."""

    def __init__(self, index: Index, max_batch_queries: int = 0, max_inflight: int = 0,
                 hbm_budget_bytes: int = 0, batch_window_us: int = 200):
        L = lib()
        L.sc_server_create.argtypes = [C.c_void_p, C.POINTER(sc_server_config),
                                       C.POINTER(C.c_void_p)]
        L.sc_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_int,
                                       C.c_uint16, C.c_int, C.POINTER(C.c_uint64)]
        L.sc_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double)]
        L.sc_server_stats.argtypes = [C.c_void_p, C.POINTER(sc_server_info)]
        L.sc_server_destroy.argtypes = [C.c_void_p]
        self.index = index
        cfg = sc_server_config(max_batch_queries, max_inflight, hbm_budget_bytes, batch_window_us)
        h = C.c_void_p()
        check(L.sc_server_create(index.handle, C.byref(cfg), C.byref(h)))
        self._h = h
        self._meta = {}

    def submit(self, queries: np.ndarray, k: int, metric: int = SC_EUCLIDIAN, p: int = 2,
               mode: int = SC_PLAIN_FP64) -> int:
        queries = np.ascontiguousarray(queries, dtype=self.index.dtype)
        if queries.ndim == 1:
            queries = queries[None, :]
        t = C.c_uint64()
        check(lib().sc_server_submit(self._h, _ptr(queries), queries.shape[0], k, metric, p, mode,
                                     C.byref(t)))
        self._meta[t.value] = (queries.shape[0], k)
        return t.value

    def wait(self, ticket: int):
        q, k = self._meta.pop(ticket)
        idx = np.empty((q, k), np.int64)
        dist = np.empty((q, k), np.float64)
        cnt = np.empty(q, np.uint32)
        lat = C.c_double()
        check(lib().sc_server_wait(self._h, ticket, _ptr(idx), _ptr(dist), _ptr(cnt),
                                   C.byref(lat)))
        return idx, dist, cnt, lat.value

    def stats(self) -> dict:
        info = sc_server_info()
        lib().sc_server_stats(self._h, C.byref(info))
        return {f: getattr(info, f) for f, _ in sc_server_info._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().sc_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class sc_kmeans_problem(C.Structure):
    _fields_ = [("rows", C.c_void_p), ("n", C.c_size_t), ("d", C.c_size_t), ("k", C.c_size_t),
                ("init_centroids", C.c_void_p)]


def kmeans_batch_run(problems, max_iter: int, mode: int = SC_MIXED):
    """sc_kmeans_batch_run over [(rows, init_centroids), ...] (all rows of one
    dtype).  Returns a list of (centroids, labels, size, distortion, iters);
    the arrays of all problems are views of three result buffers."""
    keep = []
    arr = (sc_kmeans_problem * max(len(problems), 1))()
    dtype = None
    for i, (x, init) in enumerate(problems):
        x = np.ascontiguousarray(x)
        init = np.ascontiguousarray(init, dtype=np.float64)
        dtype = x.dtype if dtype is None else dtype
        if x.dtype != dtype:
            raise TypeError("all problems must share one dtype")
        keep += [x, init]
        arr[i] = sc_kmeans_problem(x.ctypes.data, x.shape[0], x.shape[1], init.shape[0],
                                   init.ctypes.data)
    nk = sum(init.shape[0] * init.shape[1] for _, init in problems)
    nl = sum(x.shape[0] for x, _ in problems)
    ns = sum(init.shape[0] for _, init in problems)
    cent = np.empty(max(nk, 1))
    labels = np.empty(max(nl, 1), np.uint32)
    sizes = np.empty(max(ns, 1), np.uint64)
    dist = np.empty(max(len(problems), 1))
    iters = np.empty(max(len(problems), 1), np.uint32)
    lib().sc_kmeans_batch_run.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_size_t, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
    check(lib().sc_kmeans_batch_run(C.cast(arr, C.c_void_p), len(problems),
                                    dtype_code(dtype) if dtype is not None else SC_F32, max_iter,
                                    mode, _ptr(cent), _ptr(labels), _ptr(sizes), _ptr(dist),
                                    _ptr(iters)))
    out = []
    co = lo = so = 0
    for i, (x, init) in enumerate(problems):
        k, d = init.shape
        n = x.shape[0]
        out.append((cent[co:co + k * d].reshape(k, d), labels[lo:lo + n], sizes[so:so + k],
                    float(dist[i]), int(iters[i])))
        co += k * d
        lo += n
        so += k
    return out


def kmeans_predict(x: np.ndarray, centroids: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    centroids = np.ascontiguousarray(centroids, dtype=np.float64)
    n, d = x.shape
    labels = np.zeros(max(n, 1), np.uint32)
    check(lib().sc_kmeans_predict(_ptr(x), dtype_code(x.dtype), n, d, _ptr(centroids),
                                  centroids.shape[0], _ptr(labels)))
    return labels[:n]
