/*
 * smartcore_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * smartcore's distance-bound hot path: brute-force k-nearest-neighbour
 * search and the k-means Lloyd step.
 *
 * This is the boundary a smartcore build would bind through `extern "C"`
 * declarations in a Rust stub.  Every
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
 * distances, heap-replayed output).  SC_MIXED: for Euclidian, BF16 MFMA
 * cross terms with FP32 accumulation select a certified shortlist; for
 * Manhattan, FP32 sums with a rigorous error bound do.  The shortlist is
 * rescored in the SC_PLAIN_FP64 arithmetic, so the output is identical to
 * SC_PLAIN_FP64's (queries the certificate cannot cover are recomputed in
 * SC_PLAIN_FP64).  Minkowski has no SC_MIXED form. */
typedef enum { SC_PLAIN_FP64 = 0, SC_MIXED = 1 } sc_mode;

/* Per-call counters of the calling thread's last search (sc_last_stats). */
typedef struct {
  uint64_t queries;            /* queries searched */
  uint64_t candidates;         /* (index, distance) candidates replayed */
  uint64_t overflow_queries;   /* queries recomputed by the sequential kernel */
  uint64_t shortlist_rescored; /* SC_MIXED: shortlist rows rescored in FP64 */
  uint64_t uncertified;        /* SC_MIXED: queries whose shortlist was not certified */
  uint64_t launch_waves;       /* batched k-means: waves per workgroup the launch-shape
                                  tuner chose (0 for other calls) */
  uint64_t launch_groups_per_cu; /* batched k-means: workgroups per CU it chose */
  uint64_t upload_groups;      /* batched k-means: groups of problems whose rows were
                                  copied while the previous group iterated */
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

/* Many uneven query blocks against one index in one call (no reference
 * counterpart: smartcore calls LinearKNNSearch::find, linear_search.rs:52-84,
 * once per row; this is the batched form of that loop for many independent
 * callers).  Each block has its own host queries (q*d, index dtype), k and
 * outputs in the sc_knn_search layout.  Blocks with the same k run as one
 * device search over their concatenated rows; every block gets exactly what
 * sc_knn_search gives it alone.  Errors as sc_knn_search, for the first
 * block that fails validation. */
typedef struct {
  const void* queries;
  size_t q;
  size_t k;
  int64_t* idx;
  double* dist;
  uint32_t* count;
} sc_knn_block;
int sc_knn_search_batch(const sc_index* index, const sc_knn_block* blocks, size_t count,
                        sc_metric metric, uint16_t p, sc_mode mode);

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
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sc_kmeans_step with device pointers (d_centroids k*d
 * f64; d_sums k*d, d_counts k, d_membership n, d_distortion 1), ordered on
 * `stream`, returning after enqueueing.  It allocates nothing: its scratch
 * is one fixed 64 MiB device array per GPU, held until the process ends.
 * Shapes whose scratch exceeds it (about 8*(nb+ng)*k*d + 12*nb*k + 8*k*ld
 * bytes, nb = ceil(n/2048), ng = ceil(nb/16)) fail with SC_ERR_UNSUPPORTED.
 * All calls on one GPU share that array: calls that overlap on one GPU
 * race, so the caller serialises them.  Outputs equal sc_kmeans_step's bit
 * for bit. */
int sc_kmeans_step_device_noalloc(const sc_index* index, const double* d_centroids, size_t k,
                                  double* d_sums, uint64_t* d_counts, uint32_t* d_membership,
                                  double* d_distortion, void* stream);
/* The Lloyd loop of KMeans::fit (kmeans.rs:294-310) from caller-supplied
 * initial centroids (k*d f64, updated in place): at most max_iter steps,
 * centroid_i = sums_i / size_i when size_i > 0, stop when the previous
 * distortion <= the new one.  mode SC_MIXED (f32 rows, k <= 128, d <= 128,
 * not both k > 64 and d > 64) runs the assignment on BF16 MFMA cross terms certified against the FP64
 * rule (rows the bound cannot separate are reassigned in FP64) fused with the
 * FP64 sums, so its labels, sizes, centroids, distortion and iteration count
 * equal SC_PLAIN_FP64's bit for bit; other shapes run SC_PLAIN_FP64.  Errors:
 * SC_FIT_FAILED "invalid number of clusters: k" / "invalid maximum number of
 * iterations: m" (kmeans.rs:257-269). */
int sc_kmeans_run(const sc_index* index, size_t k, size_t max_iter,
                  sc_mode mode, double* centroids, uint32_t* labels,
                  uint64_t* size, double* distortion, uint32_t* iters);
/* Many independent KMeans fits at once (no reference counterpart: smartcore
 * fits one KMeans at a time, kmeans.rs:254-323; this batches that loop).
 * Each problem gives host rows (n*d row-major, `dtype`), k and k*d initial
 * centroids.  Outputs are concatenated in problem order: centroids
 * (sum k_i*d_i), labels (sum n_i), sizes (sum k_i), one distortion and one
 * iteration count per problem.  SC_MIXED with f32 rows runs every problem's
 * certified MFMA Lloyd step in one persistent launch per shape class per
 * iteration, followed by two reduce launches over all problems; the
 * iterations after the grid-size trials are replayed from one hipGraph.
 * Each problem's results are bit-identical to sc_kmeans_run on it alone.
 * Other modes and shapes outside the batched kernel's limits (k <= 128,
 * d <= 128, not both k > 64 and d > 64) run the problems one by one. */
typedef struct {
  const void* rows;
  size_t n;
  size_t d;
  size_t k;
  const double* init_centroids;
} sc_kmeans_problem;
int sc_kmeans_batch_run(const sc_kmeans_problem* problems, size_t count, sc_dtype dtype,
                        size_t max_iter, sc_mode mode, double* centroids_out,
                        uint32_t* labels_out, uint64_t* sizes_out, double* distortion_out,
                        uint32_t* iters_out);
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

/* ---- DBSCAN (src/cluster/dbscan.rs) ---- */
/* DBSCAN::fit (dbscan.rs:274-354) over the index's rows with find_radius
 * (linear_search.rs:89-108) as its neighbourhood query: labels (n) get -1
 * for outliers and the cluster id otherwise, num_classes the number of
 * clusters.  The neighbourhoods of all rows are computed on the device in
 * one batched radius search; the expansion loop runs on the host in the
 * reference's order.  Errors: SC_FIT_FAILED "Invalid minPts"
 * (min_samples < 1), "Invalid radius: " (eps <= 0). */
int sc_dbscan_fit(const sc_index* index, double eps, size_t min_samples, sc_metric metric,
                  uint16_t p, int32_t* labels, size_t* num_classes);
/* DBSCAN::predict (dbscan.rs:358-397): for q host query rows (row-major,
 * index dtype), the first most frequent fitted label among the index rows
 * within eps (outliers pooled last); out gets that label + 1, or 0 when the
 * outliers win.  Error: SC_FIND_FAILED "radius should be > 0". */
int sc_dbscan_predict(const sc_index* index, const int32_t* labels, size_t num_classes,
                      double eps, sc_metric metric, uint16_t p, const void* queries, size_t q,
                      int32_t* out);

/* ---- kNN serving loop (no reference counterpart: smartcore answers one
 * find() at a time; this serves LinearKNNSearch::find, linear_search.rs:52-84,
 * to concurrent callers) ----
 * One index stays resident.  sc_server_submit queues a request of q query
 * rows (host memory, copied) and returns a ticket at once; sc_server_wait
 * blocks until that request's rows are answered (same layout as
 * sc_knn_search) and gives its submit-to-result latency.  A scheduler thread
 * admits queued requests while the HBM their batch needs fits the budget,
 * coalesces requests with the same (metric, p, mode, k) into batches of up
 * to max_batch_queries rows, and keeps up to max_inflight batches running
 * concurrently on their own streams (continuous batching).  Results do not
 * depend on which requests share a batch. */
typedef struct sc_server sc_server;
typedef struct {
  size_t max_batch_queries; /* rows per device batch (0: 4096) */
  size_t max_inflight;      /* batches on the device at once (0: 3) */
  size_t hbm_budget_bytes;  /* admission budget (0: 80% of free HBM) */
  uint32_t batch_window_us; /* wait for company when a lone small request is queued */
} sc_server_config;
typedef struct {
  uint64_t completed;  /* requests answered */
  uint64_t batches;    /* device batches launched */
  uint64_t refused;    /* requests larger than the whole budget */
  uint64_t queries;    /* query rows answered */
  uint64_t pending;    /* requests waiting for admission */
  uint64_t hbm_in_use; /* admitted bytes now */
  uint64_t hbm_peak;   /* largest admitted bytes */
  uint64_t hbm_budget;
} sc_server_info;
int sc_server_create(const sc_index* index, const sc_server_config* cfg, sc_server** out);
/* Errors as sc_knn_search (SC_FIND_FAILED for k outside 1..n, ...);
 * SC_ERR_OUT_OF_MEMORY when the request alone exceeds the budget. */
int sc_server_submit(sc_server* server, const void* queries, size_t q, size_t k,
                     sc_metric metric, uint16_t p, sc_mode mode, uint64_t* ticket);
int sc_server_wait(sc_server* server, uint64_t ticket, int64_t* idx, double* dist,
                   uint32_t* count, double* latency_ms);
int sc_server_stats(sc_server* server, sc_server_info* out);
/* Waits for launched batches; queued requests fail with SC_INVALID_STATE. */
int sc_server_destroy(sc_server* server);

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
