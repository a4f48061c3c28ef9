/*
 * smartcore_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * smartcore's distance-bound hot path: brute-force k-nearest-neighbour
 * search and the k-means Lloyd step.
 *
 * This is the boundary a smartcore build would bind through `extern "C"`
 * declarations (the Rust stub is written out in INTEGRATION.md).  Every
 * entry point names the reference interface it replaces (paths relative to
 * the smartcore v0.4.0 crate root).  Pointers are plain host pointers unless
 * the entry point's name ends in `_device`, in which case they are device
 * (HBM) pointers and the call is ordered on the given HIP stream.
 *
 * Status codes: SC_OK, or the discriminant of smartcore's FailedError
 * (src/error/mod.rs:20-37) for the errors the reference returns, or one of
 * the SC_ERR_* codes >= 64 for conditions the reference turns into a panic
 * or that exist only on the device.  sc_last_error() returns the message of
 * the calling thread's last failure, worded as the reference words it
 * ("k should be >= 1 and <= length(data)", ...).
 *
 * Calls are blocking and synchronous (except *_device ones, which are
 * stream-ordered), like the reference's.  Concurrent calls on one index from
 * several host threads are safe: each call owns its workspace and stream.
 */
#ifndef SMARTCORE_AMD_H
#define SMARTCORE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define SC_OK 0
/* FailedError discriminants, src/error/mod.rs:20-37 */
#define SC_FIT_FAILED 1
#define SC_PREDICT_FAILED 2
#define SC_TRANSFORM_FAILED 3
#define SC_FIND_FAILED 4
#define SC_DECOMPOSITION_FAILED 5
#define SC_SOLUTION_FAILED 6
#define SC_PARAMETERS_ERROR 7
#define SC_INVALID_STATE 8
/* conditions the reference panics on (euclidian.rs:52-54 "Input vector sizes
 * are different.", minkowski.rs:57-59 "p must be at least 1") */
#define SC_ERR_PANIC 64
/* HIP runtime failure (message carries hipGetErrorString) */
#define SC_ERR_DEVICE 65
/* device memory could not be reserved for the call */
#define SC_ERR_OUT_OF_MEMORY 66
/* argument outside what this implementation supports (message says which) */
#define SC_ERR_UNSUPPORTED 67

/* element type of the rows; smartcore's Number impls for f64/f32
 * (src/numbers/basenum.rs:26-36) */
typedef enum { SC_F64 = 0, SC_F32 = 1 } sc_dtype;

/* the three Distance impls on the path: Euclidian (euclidian.rs:51-72),
 * Manhattan (manhattan.rs:50-62), Minkowski (minkowski.rs:50-67) */
typedef enum { SC_EUCLIDIAN = 0, SC_MANHATTAN = 1, SC_MINKOWSKI = 2 } sc_metric;

/* SC_PLAIN_FP64: distances in the reference's own arithmetic order (bit-exact
 * distances, heap-replayed output).  SC_MIXED: Euclidian only; BF16 MFMA
 * cross terms with FP32 accumulation select a certified shortlist that is
 * rescored in the SC_PLAIN_FP64 arithmetic, so the output is identical to
 * SC_PLAIN_FP64's (queries the certificate cannot cover are recomputed in
 * SC_PLAIN_FP64). */
typedef enum { SC_PLAIN_FP64 = 0, SC_MIXED = 1 } sc_mode;

/* Per-call counters of the calling thread's last search (sc_last_stats). */
typedef struct {
  uint64_t queries;            /* queries searched */
  uint64_t candidates;         /* (index, distance) candidates replayed */
  uint64_t overflow_queries;   /* queries recomputed by the sequential kernel */
  uint64_t shortlist_rescored; /* SC_MIXED: shortlist rows rescored in FP64 */
  uint64_t uncertified;        /* SC_MIXED: queries whose shortlist was not certified */
} sc_stats;

/* Opaque device-resident reference set (packed row-major copy in HBM).
 * Replaces the Vec<Vec<T>> that LinearKNNSearch owns
 * (src/algorithm/neighbour/linear_search.rs:36-39, new :45-47) and the data
 * BBDTree is built over (bbd_tree.rs:42-60). */
typedef struct sc_index sc_index;

/* LinearKNNSearch::new (linear_search.rs:45-47) /
 * KNNAlgorithmName::fit (algorithm/neighbour/mod.rs:70-84): rows is n*d
 * row-major host memory of `dtype`. */
int sc_index_create(const void* rows, sc_dtype dtype, size_t n, size_t d,
                    sc_index** out);
/* Same, from rows already in HBM (copied; the caller keeps ownership). */
int sc_index_create_device(const void* d_rows, sc_dtype dtype, size_t n,
                           size_t d, sc_index** out);
/* Drop of LinearKNNSearch / BBDTree. */
int sc_index_destroy(sc_index* index);
size_t sc_index_rows(const sc_index* index);
size_t sc_index_dim(const sc_index* index);

/* LinearKNNSearch::find (linear_search.rs:52-84) for q query rows at once
 * (the loop of KNNClassifier::predict, knn_classifier.rs:265-277).
 * queries: q*d row-major host memory of the index's dtype.  For each query
 * i, idx[i*k .. i*k+count[i]) and dist[...] receive the neighbours in the
 * reference's heap-array order (entries whose index stayed None are dropped
 * and the row is padded with -1 / NaN).  p is Minkowski's order (ignored by
 * the other metrics).  Errors: SC_FIND_FAILED "k should be >= 1 and <=
 * length(data)"; SC_ERR_PANIC for Minkowski p < 1. */
int sc_knn_search(const sc_index* index, const void* queries, size_t q,
                  size_t k, sc_metric metric, uint16_t p, sc_mode mode,
                  int64_t* idx, double* dist, uint32_t* count);
/* Same with queries and outputs in HBM, ordered on `stream` (hipStream_t,
 * NULL = the default stream).  Returns after enqueueing unless an error
 * needs host work; the outputs are valid once the stream reaches the point
 * of the call. */
int sc_knn_search_device(const sc_index* index, const void* d_queries,
                         size_t q, size_t k, sc_metric metric, uint16_t p,
                         sc_mode mode, int64_t* d_idx, double* d_dist,
                         uint32_t* d_count, void* stream);

/* LinearKNNSearch::find_radius (linear_search.rs:89-108) for q queries.
 * Result rows are in index order.  Errors: SC_FIND_FAILED "radius should be
 * > 0". */
typedef struct sc_radius_result sc_radius_result;
int sc_radius_search(const sc_index* index, const void* queries, size_t q,
                     double radius, sc_metric metric, uint16_t p,
                     sc_radius_result** out);
/* total number of (query, neighbour) pairs */
size_t sc_radius_result_total(const sc_radius_result* r);
/* offsets: q+1 entries (CSR row starts), idx/dist: total entries */
int sc_radius_result_copy(const sc_radius_result* r, uint64_t* offsets,
                          int64_t* idx, double* dist);
void sc_radius_result_free(sc_radius_result* r);

/* BBDTree::clustering (bbd_tree.rs:62-87): one Lloyd step over the index's
 * rows.  centroids: k*d f64.  Writes sums (k*d), counts (k), membership (n)
 * and the distortion.  The assignment is KMeans::predict's rule
 * (kmeans.rs:333-347: squared distance of the widened row, strict <, lowest
 * index wins), which the filtering recursion reproduces off exact ties. */
int sc_kmeans_step(const sc_index* index, const double* centroids, size_t k,
                   double* sums, uint64_t* counts, uint32_t* membership,
                   double* distortion);
/* The Lloyd loop of KMeans::fit (kmeans.rs:294-310) from caller-supplied
 * initial centroids (k*d f64, updated in place): at most max_iter steps,
 * centroid_i = sums_i / size_i when size_i > 0, stop when the previous
 * distortion <= the new one.  mode SC_MIXED runs the assignment on BF16x3
 * MFMA cross terms certified against the FP64 rule (uncertain rows are
 * reassigned in FP64), so labels equal SC_PLAIN_FP64's.  Errors:
 * SC_FIT_FAILED "invalid number of clusters: k" / "invalid maximum number of
 * iterations: m" (kmeans.rs:257-269). */
int sc_kmeans_run(const sc_index* index, size_t k, size_t max_iter,
                  sc_mode mode, double* centroids, uint32_t* labels,
                  uint64_t* size, double* distortion, uint32_t* iters);
/* KMeans::kmeans_plus_plus (kmeans.rs:354-413): initial labels (n) of the
 * index's rows for k clusters.  The random stream is rand 0.8.5's SmallRng
 * seeded by get_rng_impl (rand_custom.rs:8-33: has_seed ? seed : 0).  The
 * per-round distance updates run on the device in the rows' own type; the
 * sequential cost sum and cumulative scan run on the host in row order. */
int sc_kmeans_plus_plus(const sc_index* index, size_t k, int has_seed, uint64_t seed,
                        uint32_t* labels);
/* KMeans::predict (kmeans.rs:327-352): labels of n rows of x (host,
 * row-major, dtype) against k centroids (k*d f64). */
int sc_kmeans_predict(const void* x, sc_dtype dtype, size_t n, size_t d,
                      const double* centroids, size_t k, uint32_t* labels);

/* Message of the calling thread's last failure; returns its full length
 * (writes at most len-1 bytes plus a terminating NUL). */
size_t sc_last_error(char* buf, size_t len);
/* Counters of the calling thread's last sc_knn_search* call. */
int sc_last_stats(sc_stats* out);

/* Kernel tracing.  The reference has no tracing of its own (profiling is
 * done out of tree, .gitignore:22-24); these hooks time the device kernels.
 * While enabled, every launch of a kernel family is bracketed by a pair of
 * HIP events on the stream it runs on.  sc_profile_enable(1) clears earlier
 * records.  sc_profile_read waits for the recorded events and returns the
 * summed duration and the launch count of one family: "knn_scan",
 * "knn_replay", "knn_serial", "knn_mixed_gemm", "knn_mixed_rescore",
 * "kmeans_assign", "kmeans_partial", "kmeans_reduce", "kmeans_mixed_assign". */
int sc_profile_enable(int on);
int sc_profile_read(const char* kernel, double* total_ms, uint64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
