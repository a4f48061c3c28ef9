/*
 * oracle.c -- CPU restatement of smartcore v0.4.0's distance-bound path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and single-core CPU baseline);
 * the product library never links or calls it.  Build with
 * -O2 -ffp-contract=off and without -ffast-math so that every sum is the
 * reference's sequential left fold with separate roundings.
 *
 * Citations are smartcore paths relative to the crate root.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ------------------------------------------------------------------ */
/* distances                                                            */
/* ------------------------------------------------------------------ */

/* src/metrics/distance/euclidian.rs:51-64 for T = f64:
 * r = a - b; sum += (r * r) as f64, left fold from 0. */
double or_squared_distance_f64(const double* a, const double* b, size_t d) {
  double sum = 0.0;
  for (size_t i = 0; i < d; ++i) {
    double r = a[i] - b[i];
    sum += r * r;
  }
  return sum;
}

/* euclidian.rs:51-64 for T = f32: the difference and the square are
 * rounded in f32 (`a - b` and `r * r` are f32 ops) and only the square is
 * widened with to_f64() before the f64 fold. */
double or_squared_distance_f32(const float* a, const float* b, size_t d) {
  double sum = 0.0;
  for (size_t i = 0; i < d; ++i) {
    float r = a[i] - b[i];
    float r2 = r * r;
    sum += (double)r2;
  }
  return sum;
}

static double elem_f(int dtype, const void* v, size_t i) {
  return dtype == OR_F64 ? ((const double*)v)[i] : (double)((const float*)v)[i];
}

/* difference in T, widened: (a - b).to_f64() */
static double diff_t(int dtype, const void* a, const void* b, size_t i) {
  if (dtype == OR_F64) return ((const double*)a)[i] - ((const double*)b)[i];
  float r = ((const float*)a)[i] - ((const float*)b)[i];
  return (double)r;
}

double or_distance(int metric, unsigned p, int dtype, const void* a,
                   const void* b, size_t d, int* err) {
  if (err) *err = 0;
  switch (metric) {
    case OR_EUCLIDIAN: {
      /* euclidian.rs:69-72: squared_distance(x, y).sqrt() */
      double s = dtype == OR_F64
                     ? or_squared_distance_f64((const double*)a, (const double*)b, d)
                     : or_squared_distance_f32((const float*)a, (const float*)b, d);
      return sqrt(s);
    }
    case OR_MANHATTAN: {
      /* manhattan.rs:50-62: sum of (a - b).to_f64().abs() */
      double s = 0.0;
      for (size_t i = 0; i < d; ++i) s += fabs(diff_t(dtype, a, b, i));
      return s;
    }
    case OR_MINKOWSKI: {
      /* minkowski.rs:50-67: p < 1 panics; sum of |a-b|^p then ^(1/p) */
      if (p < 1) {
        if (err) *err = 1;
        return NAN;
      }
      double pt = (double)p;
      double s = 0.0;
      for (size_t i = 0; i < d; ++i) s += pow(fabs(diff_t(dtype, a, b, i)), pt);
      return pow(s, 1.0 / pt);
    }
  }
  if (err) *err = 2;
  return NAN;
}

/* ------------------------------------------------------------------ */
/* HeapSelection                                                        */
/* ------------------------------------------------------------------ */

/* heap_select.rs:71-86.  Children of kk are 2*kk and 2*kk+1, so node 0's
 * only distinct child is node 1.  Comparisons are on distance only
 * (KNNPoint's PartialOrd, linear_search.rs:117-121). */
void or_sift_down(or_knn_point* heap, size_t k, size_t n) {
  size_t kk = k;
  while (2 * kk <= n) {
    size_t j = 2 * kk;
    if (j < n && heap[j].distance < heap[j + 1].distance) j += 1;
    /* break on Equal or Greater; a NaN compares as neither and does not
     * break, exactly like partial_cmp returning None. */
    if (heap[kk].distance == heap[j].distance ||
        heap[kk].distance > heap[j].distance)
      break;
    or_knn_point t = heap[kk];
    heap[kk] = heap[j];
    heap[j] = t;
    kk = j;
  }
}

void or_heap_init(or_heap* h, or_knn_point* storage, size_t k) {
  /* heap_select.rs:16-23 */
  h->k = k;
  h->n = 0;
  h->sorted = 0;
  h->len = 0;
  h->heap = storage;
}

/* heap_select.rs:88-91: sort_by(|a, b| b.partial_cmp(a).unwrap()) is a
 * stable descending sort; insertion sort is stable. */
static void heap_sort_desc(or_heap* h) {
  h->sorted = 1;
  for (size_t i = 1; i < h->len; ++i) {
    or_knn_point x = h->heap[i];
    size_t j = i;
    while (j > 0 && h->heap[j - 1].distance < x.distance) {
      h->heap[j] = h->heap[j - 1];
      --j;
    }
    h->heap[j] = x;
  }
}

/* heap_select.rs:25-40 */
void or_heap_add(or_heap* h, double distance, int64_t index) {
  h->sorted = 0;
  if (h->n < h->k) {
    h->heap[h->len].distance = distance;
    h->heap[h->len].index = index;
    h->len += 1;
    h->n += 1;
    if (h->n == h->k) heap_sort_desc(h);
  } else {
    h->n += 1;
    if (distance < h->heap[0].distance) {
      h->heap[0].distance = distance;
      h->heap[0].index = index;
      or_sift_down(h->heap, 0, h->k - 1);
    }
  }
}

/* heap_select.rs:42-50 */
void or_heap_heapify(or_heap* h) {
  size_t n = h->len;
  if (n <= 1) return;
  for (size_t i = n / 2; i-- > 0;) or_sift_down(h->heap, i, n - 1);
}

/* ------------------------------------------------------------------ */
/* LinearKNNSearch                                                      */
/* ------------------------------------------------------------------ */

static const void* row_ptr(const void* data, int dtype, size_t d, size_t i) {
  return dtype == OR_F64 ? (const void*)((const double*)data + i * d)
                         : (const void*)((const float*)data + i * d);
}

static void collect(const or_heap* h, int64_t* out_idx, double* out_dist,
                    size_t* out_count) {
  /* linear_search.rs:79-83: heap-array order, None entries dropped */
  size_t c = 0;
  for (size_t j = 0; j < h->len; ++j) {
    if (h->heap[j].index < 0) continue;
    out_idx[c] = h->heap[j].index;
    out_dist[c] = h->heap[j].distance;
    ++c;
  }
  *out_count = c;
}

/* linear_search.rs:52-84 */
int or_linear_find(const void* data, int dtype, size_t n, size_t d,
                   const void* query, size_t k, int metric, unsigned p,
                   int64_t* out_idx, double* out_dist, size_t* out_count) {
  *out_count = 0;
  if (k < 1 || k > n) return 4; /* FindFailed */
  if (metric == OR_MINKOWSKI && p < 1) return 7;
  or_knn_point* st = (or_knn_point*)malloc(k * sizeof(or_knn_point));
  or_heap h;
  or_heap_init(&h, st, k);
  for (size_t i = 0; i < k; ++i) or_heap_add(&h, INFINITY, -1);
  for (size_t i = 0; i < n; ++i) {
    double dd = or_distance(metric, p, dtype, query, row_ptr(data, dtype, d, i), d, NULL);
    or_knn_point* datum = &h.heap[0]; /* peek_mut, heap_select.rs:63-65 */
    if (dd < datum->distance) {
      datum->distance = dd;
      datum->index = (int64_t)i;
      or_heap_heapify(&h);
    }
  }
  collect(&h, out_idx, out_dist, out_count);
  free(st);
  return 0;
}

int or_linear_find_via_add(const void* data, int dtype, size_t n, size_t d,
                           const void* query, size_t k, int metric, unsigned p,
                           int64_t* out_idx, double* out_dist,
                           size_t* out_count) {
  *out_count = 0;
  if (k < 1 || k > n) return 4;
  if (metric == OR_MINKOWSKI && p < 1) return 7;
  or_knn_point* st = (or_knn_point*)malloc(k * sizeof(or_knn_point));
  or_heap h;
  or_heap_init(&h, st, k);
  for (size_t i = 0; i < k; ++i) or_heap_add(&h, INFINITY, -1);
  for (size_t i = 0; i < n; ++i) {
    double dd = or_distance(metric, p, dtype, query, row_ptr(data, dtype, d, i), d, NULL);
    or_heap_add(&h, dd, (int64_t)i);
  }
  collect(&h, out_idx, out_dist, out_count);
  free(st);
  return 0;
}

int or_linear_find_batch(const void* data, int dtype, size_t n, size_t d,
                         const void* queries, size_t q, size_t k, int metric,
                         unsigned p, int64_t* out_idx, double* out_dist,
                         uint32_t* out_count) {
  for (size_t i = 0; i < q; ++i) {
    size_t c = 0;
    int rc = or_linear_find(data, dtype, n, d, row_ptr(queries, dtype, d, i), k,
                            metric, p, out_idx + i * k, out_dist + i * k, &c);
    if (rc) return rc;
    for (size_t j = c; j < k; ++j) {
      out_idx[i * k + j] = -1;
      out_dist[i * k + j] = NAN;
    }
    out_count[i] = (uint32_t)c;
  }
  return 0;
}

/* linear_search.rs:89-108 */
int or_linear_find_radius(const void* data, int dtype, size_t n, size_t d,
                          const void* query, double radius, int metric,
                          unsigned p, int64_t* out_idx, double* out_dist,
                          size_t cap, size_t* out_count) {
  *out_count = 0;
  if (radius <= 0.0) return 4;
  if (metric == OR_MINKOWSKI && p < 1) return 7;
  size_t c = 0;
  for (size_t i = 0; i < n; ++i) {
    double dd = or_distance(metric, p, dtype, query, row_ptr(data, dtype, d, i), d, NULL);
    if (dd <= radius) {
      if (c < cap) {
        out_idx[c] = (int64_t)i;
        out_dist[c] = dd;
      }
      ++c;
    }
  }
  *out_count = c;
  return 0;
}

/* ------------------------------------------------------------------ */
/* KNN estimators                                                       */
/* ------------------------------------------------------------------ */

/* src/neighbors/mod.rs:62-78 */
void or_calc_weights(int weight, const double* dist, size_t k, double* w) {
  if (weight == 1) {
    int any_zero = 0;
    for (size_t i = 0; i < k; ++i)
      if (dist[i] == 0.0) any_zero = 1;
    for (size_t i = 0; i < k; ++i)
      w[i] = any_zero ? (dist[i] == 0.0 ? 1.0 : 0.0) : 1.0 / dist[i];
  } else {
    for (size_t i = 0; i < k; ++i) w[i] = 1.0;
  }
}

/* knn_classifier.rs:279-299 */
size_t or_knn_classifier_predict_row(const int64_t* idx, const double* dist,
                                     size_t count, const int64_t* y_class,
                                     size_t n_classes, int weight) {
  double* w = (double*)malloc((count ? count : 1) * sizeof(double));
  double* c = (double*)calloc(n_classes ? n_classes : 1, sizeof(double));
  or_calc_weights(weight, dist, count, w);
  double w_sum = 0.0;
  for (size_t i = 0; i < count; ++i) w_sum += w[i];
  double max_c = 0.0;
  size_t max_i = 0;
  for (size_t i = 0; i < count; ++i) {
    size_t cls = (size_t)y_class[idx[i]];
    c[cls] += w[i] / w_sum;
    if (c[cls] > max_c) {
      max_c = c[cls];
      max_i = cls;
    }
  }
  free(w);
  free(c);
  return max_i;
}

/* knn_regressor.rs:268-281 for TY = f64 */
double or_knn_regressor_predict_row(const int64_t* idx, const double* dist,
                                    size_t count, const double* y, int weight) {
  double* w = (double*)malloc((count ? count : 1) * sizeof(double));
  or_calc_weights(weight, dist, count, w);
  double w_sum = 0.0;
  for (size_t i = 0; i < count; ++i) w_sum += w[i];
  double result = 0.0;
  for (size_t i = 0; i < count; ++i) result += y[idx[i]] * (w[i] / w_sum);
  free(w);
  return result;
}

/* ------------------------------------------------------------------ */
/* BBDTree                                                              */
/* ------------------------------------------------------------------ */

struct or_bbd {
  size_t n, d;
  size_t nnodes, cap;
  /* per node */
  size_t* count;
  size_t* index;
  int64_t* lower;
  int64_t* upper;
  double* cost;
  double* center; /* nnodes x d */
  double* radius;
  double* sum;
  size_t* perm; /* BBDTree::index */
  size_t root;
  const void* data;
  int dtype;
};

static inline double get_f(const or_bbd* t, size_t i, size_t j) {
  return elem_f(t->dtype, t->data, i * t->d + j);
}

/* bbd_tree.rs:297-305 */
static double node_cost(const or_bbd* t, size_t node, const double* center) {
  size_t d = t->d;
  double scatter = 0.0;
  const double* s = t->sum + node * d;
  for (size_t i = 0; i < d; ++i) {
    double x = (s[i] / (double)t->count[node]) - center[i];
    scatter += x * x;
  }
  return t->cost[node] + (double)t->count[node] * scatter;
}

/* bbd_tree.rs:195-295.  Nodes are appended after their children
 * (add_node happens on return), matching the reference's indices. */
static size_t build_node(or_bbd* t, size_t begin, size_t end, double* scratch) {
  size_t d = t->d;
  (void)scratch;
  double* buf = (double*)malloc(5 * (d ? d : 1) * sizeof(double));
  double* lb = buf;
  double* ub = buf + d;
  double* center = buf + 2 * d;
  double* radius = buf + 3 * d;
  double* mean = buf + 4 * d;
  for (size_t i = 0; i < d; ++i) {
    lb[i] = get_f(t, t->perm[begin], i);
    ub[i] = get_f(t, t->perm[begin], i);
  }
  for (size_t i = begin; i < end; ++i) {
    for (size_t j = 0; j < d; ++j) {
      double c = get_f(t, t->perm[i], j);
      if (lb[j] > c) lb[j] = c;
      if (ub[j] < c) ub[j] = c;
    }
  }
  double max_radius = -1.0;
  size_t split_index = 0;
  for (size_t i = 0; i < d; ++i) {
    center[i] = (lb[i] + ub[i]) / 2.0;
    radius[i] = (ub[i] - lb[i]) / 2.0;
    if (radius[i] > max_radius) {
      max_radius = radius[i];
      split_index = i;
    }
  }
  size_t count = end - begin;
  if (max_radius < 1e-10) {
    size_t id = t->nnodes++;
    t->count[id] = count;
    t->index[id] = begin;
    t->lower[id] = -1;
    t->upper[id] = -1;
    memcpy(t->center + id * d, center, d * sizeof(double));
    memcpy(t->radius + id * d, radius, d * sizeof(double));
    double* s = t->sum + id * d;
    for (size_t i = 0; i < d; ++i) s[i] = get_f(t, t->perm[begin], i);
    if (end > begin + 1) {
      for (size_t i = 0; i < d; ++i) s[i] *= (double)count;
    }
    t->cost[id] = 0.0;
    free(buf);
    return id;
  }
  double split_cutoff = center[split_index];
  int64_t i1 = (int64_t)begin;
  int64_t i2 = (int64_t)end - 1;
  size_t size = 0;
  while (i1 <= i2) {
    int i1_good = get_f(t, t->perm[i1], split_index) < split_cutoff;
    int i2_good = get_f(t, t->perm[i2], split_index) >= split_cutoff;
    if (!i1_good && !i2_good) {
      size_t tmp = t->perm[i1];
      t->perm[i1] = t->perm[i2];
      t->perm[i2] = tmp;
      i1_good = 1;
      i2_good = 1;
    }
    if (i1_good) {
      i1 += 1;
      size += 1;
    }
    if (i2_good) i2 -= 1;
  }
  size_t lower = build_node(t, begin, begin + size, scratch);
  size_t upper = build_node(t, begin + size, end, scratch);
  size_t id = t->nnodes++;
  t->count[id] = count;
  t->index[id] = begin;
  t->lower[id] = (int64_t)lower;
  t->upper[id] = (int64_t)upper;
  memcpy(t->center + id * d, center, d * sizeof(double));
  memcpy(t->radius + id * d, radius, d * sizeof(double));
  double* s = t->sum + id * d;
  for (size_t i = 0; i < d; ++i) s[i] = t->sum[lower * d + i] + t->sum[upper * d + i];
  for (size_t i = 0; i < d; ++i) mean[i] = s[i] / (double)count;
  t->cost[id] = node_cost(t, lower, mean) + node_cost(t, upper, mean);
  free(buf);
  return id;
}

/* bbd_tree.rs:42-60 */
or_bbd* or_bbd_new(const void* data, int dtype, size_t n, size_t d) {
  or_bbd* t = (or_bbd*)calloc(1, sizeof(or_bbd));
  t->n = n;
  t->d = d;
  t->data = data;
  t->dtype = dtype;
  t->cap = 2 * n + 1;
  t->count = (size_t*)malloc(t->cap * sizeof(size_t));
  t->index = (size_t*)malloc(t->cap * sizeof(size_t));
  t->lower = (int64_t*)malloc(t->cap * sizeof(int64_t));
  t->upper = (int64_t*)malloc(t->cap * sizeof(int64_t));
  t->cost = (double*)malloc(t->cap * sizeof(double));
  t->center = (double*)malloc(t->cap * d * sizeof(double));
  t->radius = (double*)malloc(t->cap * d * sizeof(double));
  t->sum = (double*)malloc(t->cap * d * sizeof(double));
  t->perm = (size_t*)malloc((n ? n : 1) * sizeof(size_t));
  for (size_t i = 0; i < n; ++i) t->perm[i] = i;
  double* scratch = (double*)malloc(2 * (d ? d : 1) * sizeof(double));
  t->root = build_node(t, 0, n, scratch);
  free(scratch);
  return t;
}

void or_bbd_free(or_bbd* t) {
  if (!t) return;
  free(t->count);
  free(t->index);
  free(t->lower);
  free(t->upper);
  free(t->cost);
  free(t->center);
  free(t->radius);
  free(t->sum);
  free(t->perm);
  free(t);
}

size_t or_bbd_num_nodes(const or_bbd* t) { return t->nnodes; }

/* bbd_tree.rs:165-193 */
static int prune(const double* center, const double* radius,
                 const double* centroids, size_t d, size_t best_index,
                 size_t test_index) {
  if (best_index == test_index) return 0;
  const double* best = centroids + best_index * d;
  const double* test = centroids + test_index * d;
  double lhs = 0.0, rhs = 0.0;
  for (size_t i = 0; i < d; ++i) {
    double diff = test[i] - best[i];
    lhs += diff * diff;
    if (diff > 0.0)
      rhs += (center[i] + radius[i] - best[i]) * diff;
    else
      rhs += (center[i] - radius[i] - best[i]) * diff;
  }
  return lhs >= 2.0 * rhs;
}

/* bbd_tree.rs:89-163 */
static double filter(const or_bbd* t, size_t node, const double* centroids,
                     const size_t* candidates, size_t k, double* sums,
                     uint64_t* counts, uint32_t* membership) {
  size_t d = t->d;
  const double* ctr = t->center + node * d;
  double min_dist = or_squared_distance_f64(ctr, centroids + candidates[0] * d, d);
  size_t closest = candidates[0];
  for (size_t i = 1; i < k; ++i) {
    double dist = or_squared_distance_f64(ctr, centroids + candidates[i] * d, d);
    if (dist < min_dist) {
      min_dist = dist;
      closest = candidates[i];
    }
  }
  if (t->lower[node] >= 0) {
    size_t* nc = (size_t*)malloc(k * sizeof(size_t));
    size_t newk = 0;
    for (size_t c = 0; c < k; ++c) {
      if (!prune(ctr, t->radius + node * d, centroids, d, closest, candidates[c]))
        nc[newk++] = candidates[c];
    }
    if (newk > 1) {
      double r = filter(t, (size_t)t->lower[node], centroids, nc, newk, sums, counts, membership);
      r = r + filter(t, (size_t)t->upper[node], centroids, nc, newk, sums, counts, membership);
      free(nc);
      return r;
    }
    free(nc);
  }
  const double* s = t->sum + node * d;
  for (size_t i = 0; i < d; ++i) sums[closest * d + i] += s[i];
  counts[closest] += t->count[node];
  size_t last = t->index[node] + t->count[node];
  for (size_t i = t->index[node]; i < last; ++i) membership[t->perm[i]] = (uint32_t)closest;
  return node_cost(t, node, centroids + closest * d);
}

/* bbd_tree.rs:62-87 */
double or_bbd_clustering(const or_bbd* t, const double* centroids, size_t k,
                         double* sums, uint64_t* counts, uint32_t* membership) {
  size_t d = t->d;
  size_t* cand = (size_t*)malloc(k * sizeof(size_t));
  for (size_t i = 0; i < k; ++i) {
    counts[i] = 0;
    cand[i] = i;
    for (size_t j = 0; j < d; ++j) sums[i * d + j] = 0.0;
  }
  double r = filter(t, t->root, centroids, cand, k, sums, counts, membership);
  free(cand);
  return r;
}

/* ------------------------------------------------------------------ */
/* KMeans                                                               */
/* ------------------------------------------------------------------ */

/* kmeans.rs:275-292 */
void or_kmeans_means_from_labels(const void* data, int dtype, size_t n,
                                 size_t d, const uint32_t* y, size_t k,
                                 double* centroids, uint64_t* size) {
  for (size_t i = 0; i < k; ++i) size[i] = 0;
  for (size_t i = 0; i < k * d; ++i) centroids[i] = 0.0;
  for (size_t i = 0; i < n; ++i) size[y[i]] += 1;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < d; ++j) centroids[y[i] * d + j] += elem_f(dtype, data, i * d + j);
  for (size_t i = 0; i < k; ++i)
    for (size_t j = 0; j < d; ++j) centroids[i * d + j] /= (double)size[i];
}

/* KMeans::predict assignment rule, kmeans.rs:333-347: row widened to f64,
 * squared distance in f64, strict < so the lowest index wins ties. */
static uint32_t assign_row(const void* data, int dtype, size_t d, size_t i,
                           const double* centroids, size_t k, double* row,
                           double* best_dist) {
  for (size_t j = 0; j < d; ++j) row[j] = elem_f(dtype, data, i * d + j);
  double min_dist = 1.7976931348623157e308; /* f64::MAX */
  uint32_t best = 0;
  for (size_t j = 0; j < k; ++j) {
    double dist = or_squared_distance_f64(row, centroids + j * d, d);
    if (dist < min_dist) {
      min_dist = dist;
      best = (uint32_t)j;
    }
  }
  *best_dist = min_dist;
  return best;
}

/* ------------------------------------------------------------------ */
/* random stream of KMeans::kmeans_plus_plus                            */
/* ------------------------------------------------------------------ */
/* get_rng_impl (rand_custom.rs:8-33, default features) returns rand 0.8.5's
 * SmallRng::seed_from_u64(seed), seed 0 for None.  rand 0.8.5's SmallRng
 * (64-bit targets: xoshiro256++) implements only from_seed / from_rng, so
 * seed_from_u64 is rand_core 0.6's trait default: eight PCG32 (XSH-RR)
 * outputs, each taken after advancing the LCG state, fill the 32-byte seed
 * little-endian; from_seed reads four little-endian u64 words, and an
 * all-zero seed is replaced by xoshiro's own seed_from_u64(0) (SplitMix64).
 * The rand crates are not vendored in the reference tree; this follows their
 * published source. */

/* PCG XSH-RR output of a 64-bit LCG state. */
uint32_t or_pcg32_output(uint64_t state) {
  uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
  uint32_t rot = (uint32_t)(state >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void xoshiro_from_splitmix(uint64_t state, uint64_t s[4]) {
  for (int i = 0; i < 4; ++i) {
    state += 0x9e3779b97f4a7c15ULL;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    s[i] = z ^ (z >> 31);
  }
}

void or_smallrng_seed(uint64_t seed, uint64_t s[4]) {
  const uint64_t MUL = 6364136223846793005ULL;
  const uint64_t INC = 11634580027462260723ULL;
  uint8_t bytes[32];
  uint64_t st = seed;
  for (int c = 0; c < 8; ++c) {
    st = st * MUL + INC;
    uint32_t x = or_pcg32_output(st);
    for (int b = 0; b < 4; ++b) bytes[4 * c + b] = (uint8_t)(x >> (8 * b));
  }
  int zero = 1;
  for (int i = 0; i < 32; ++i) zero &= bytes[i] == 0;
  if (zero) {
    xoshiro_from_splitmix(0, s);
    return;
  }
  for (int w = 0; w < 4; ++w) {
    uint64_t v = 0;
    for (int b = 7; b >= 0; --b) v = (v << 8) | bytes[8 * w + b];
    s[w] = v;
  }
}

/* xoshiro256++ next_u64 */
uint64_t or_smallrng_next(uint64_t s[4]) {
  const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return result;
}

/* Rng::gen_range(0..n) for usize (rand 0.8.5 UniformInt::sample_single ->
 * sample_single_inclusive(0, n-1)): range = n; zone = (range << lz(range))
 * - 1; draw v until the low word of v * range is <= zone, return the high
 * word. */
uint64_t or_smallrng_gen_range(uint64_t s[4], uint64_t n) {
  const uint64_t range = n;
  if (range == 0) return or_smallrng_next(s);
  const uint64_t zone = (range << __builtin_clzll(range)) - 1;
  for (;;) {
    const uint64_t v = or_smallrng_next(s);
    const unsigned __int128 m = (unsigned __int128)v * range;
    if ((uint64_t)m <= zone) return (uint64_t)(m >> 64);
  }
}

/* Standard for f64: (next_u64 >> 11) * 2^-53 */
double or_smallrng_gen_f64(uint64_t s[4]) {
  const double scale = 1.0 / (double)((uint64_t)1 << 53);
  return scale * (double)(or_smallrng_next(s) >> 11);
}

/* KMeans::kmeans_plus_plus, kmeans.rs:354-413: labels of the rows. */
void or_kmeans_plus_plus(const void* data, int dtype, size_t n, size_t d, size_t k,
                         int has_seed, uint64_t seed, uint32_t* y) {
  uint64_t s[4];
  or_smallrng_seed(has_seed ? seed : 0, s);
  const size_t es = dtype == OR_F64 ? 8 : 4;
  const char* base = (const char*)data;
  double* dm = (double*)malloc((n ? n : 1) * sizeof(double));
  for (size_t i = 0; i < n; ++i) {
    y[i] = 0;
    dm[i] = DBL_MAX;
  }
  /* :358-362 */
  size_t c = (size_t)or_smallrng_gen_range(s, n);
  for (size_t j = 1; j < k; ++j) {
    /* :367-379: squared distance in the rows' own type */
    for (size_t i = 0; i < n; ++i) {
      double dist = dtype == OR_F64
                        ? or_squared_distance_f64((const double*)(base + i * d * es),
                                                  (const double*)(base + c * d * es), d)
                        : or_squared_distance_f32((const float*)(base + i * d * es),
                                                  (const float*)(base + c * d * es), d);
      if (dist < dm[i]) {
        dm[i] = dist;
        y[i] = (uint32_t)(j - 1);
      }
    }
    /* :381-394 */
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += dm[i];
    const double cutoff = or_smallrng_gen_f64(s) * sum;
    double cost = 0.0;
    size_t index = 0;
    while (index < n) {
      cost += dm[index];
      if (cost >= cutoff) break;
      index += 1;
    }
    /* the reference indexes row `index` (a panic when index == n, reachable
     * only with a NaN sum); stay on the last row */
    c = index < n ? index : n - 1;
  }
  /* :399-410 */
  for (size_t i = 0; i < n; ++i) {
    double dist = dtype == OR_F64
                      ? or_squared_distance_f64((const double*)(base + i * d * es),
                                                (const double*)(base + c * d * es), d)
                      : or_squared_distance_f32((const float*)(base + i * d * es),
                                                (const float*)(base + c * d * es), d);
    if (dist < dm[i]) {
      dm[i] = dist;
      y[i] = (uint32_t)(k - 1);
    }
  }
  free(dm);
}

void or_kmeans_predict(const void* data, int dtype, size_t n, size_t d,
                       const double* centroids, size_t k, uint32_t* labels) {
  double* row = (double*)malloc((d ? d : 1) * sizeof(double));
  double bd;
  for (size_t i = 0; i < n; ++i) labels[i] = assign_row(data, dtype, d, i, centroids, k, row, &bd);
  free(row);
}

double or_lloyd_step_brute(const void* data, int dtype, size_t n, size_t d,
                           const double* centroids, size_t k, double* sums,
                           uint64_t* counts, uint32_t* membership) {
  double* row = (double*)malloc((d ? d : 1) * sizeof(double));
  for (size_t i = 0; i < k; ++i) {
    counts[i] = 0;
    for (size_t j = 0; j < d; ++j) sums[i * d + j] = 0.0;
  }
  double dist = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double bd;
    uint32_t c = assign_row(data, dtype, d, i, centroids, k, row, &bd);
    membership[i] = c;
    counts[c] += 1;
    for (size_t j = 0; j < d; ++j) sums[c * d + j] += row[j];
    dist += bd;
  }
  free(row);
  return dist;
}

/* kmeans.rs:257-269 checks and the loop kmeans.rs:294-310 */
int or_kmeans_fit(const void* data, int dtype, size_t n, size_t d, size_t k,
                  size_t max_iter, int use_bbd, double* centroids,
                  uint32_t* labels, uint64_t* size, double* distortion_out,
                  uint32_t* iters) {
  if (k < 2 || max_iter == 0) return 1;
  or_bbd* t = use_bbd ? or_bbd_new(data, dtype, n, d) : NULL;
  double* sums = (double*)calloc(k * d, sizeof(double));
  double distortion = 1.7976931348623157e308;
  uint32_t it = 0;
  for (size_t iter = 1; iter <= max_iter; ++iter) {
    double dist = use_bbd ? or_bbd_clustering(t, centroids, k, sums, size, labels)
                          : or_lloyd_step_brute(data, dtype, n, d, centroids, k, sums, size, labels);
    ++it;
    for (size_t i = 0; i < k; ++i) {
      if (size[i] > 0) {
        for (size_t j = 0; j < d; ++j) centroids[i * d + j] = sums[i * d + j] / (double)size[i];
      }
    }
    if (distortion <= dist) break;
    distortion = dist;
  }
  *distortion_out = distortion;
  *iters = it;
  free(sums);
  or_bbd_free(t);
  return 0;
}

/* ------------------------------------------------------------------ */
/* DBSCAN                                                               */
/* ------------------------------------------------------------------ */

/* find_radius of row `i` into a fresh array (linear_search.rs:89-108). */
static int64_t* radius_of(const void* data, int dtype, size_t n, size_t d, const void* q,
                          double eps, int metric, unsigned p, size_t* cnt) {
  size_t c = 0;
  or_linear_find_radius(data, dtype, n, d, q, eps, metric, p, NULL, NULL, 0, &c);
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (c ? c : 1));
  double* dist = (double*)malloc(sizeof(double) * (c ? c : 1));
  size_t c2 = 0;
  or_linear_find_radius(data, dtype, n, d, q, eps, metric, p, idx, dist, c, &c2);
  free(dist);
  *cnt = c2;
  return idx;
}

/* dbscan.rs:274-354 */
int or_dbscan_fit(const void* data, int dtype, size_t n, size_t d, double eps,
                  size_t min_samples, int metric, unsigned p, int32_t* y,
                  size_t* num_classes) {
  const int32_t queued = -2, outlier = -1, undefined = -3;
  *num_classes = 0;
  if (min_samples < 1) return 1;
  if (eps <= 0.0) return 1;
  int32_t k = 0;
  for (size_t i = 0; i < n; ++i) y[i] = undefined;
  for (size_t i = 0; i < n; ++i) {
    if (y[i] != undefined) continue;
    size_t nc = 0;
    int64_t* nb = radius_of(data, dtype, n, d, row_ptr(data, dtype, d, i), eps, metric, p, &nc);
    if (nc < min_samples) {
      y[i] = outlier;
      free(nb);
      continue;
    }
    y[i] = k;
    for (size_t j = 0; j < nc; ++j)
      if (y[nb[j]] == undefined) y[nb[j]] = queued;
    /* `neighbors` is used as a stack: pop from the end, push at the end */
    size_t top = nc, cap = nc ? nc : 1;
    int64_t* st = nb;
    while (top > 0) {
      const int64_t index = st[--top];
      if (y[index] == outlier) y[index] = k;
      if (y[index] == undefined || y[index] == queued) {
        y[index] = k;
        size_t sc = 0;
        int64_t* sec = radius_of(data, dtype, n, d, row_ptr(data, dtype, d, (size_t)index), eps,
                                 metric, p, &sc);
        if (sc >= min_samples) {
          for (size_t j = 0; j < sc; ++j) {
            const int32_t label = y[sec[j]];
            if (label == undefined) y[sec[j]] = queued;
            if (label == undefined || label == outlier) {
              if (top == cap) {
                cap *= 2;
                st = (int64_t*)realloc(st, sizeof(int64_t) * cap);
              }
              st[top++] = sec[j];
            }
          }
        }
        free(sec);
      }
    }
    free(st);
    k += 1;
  }
  *num_classes = (size_t)k;
  return 0;
}

/* dbscan.rs:358-397 with which_max (decision_tree_classifier.rs:490-502):
 * the first maximum wins. */
int or_dbscan_predict(const void* data, int dtype, size_t n, size_t d,
                      const int32_t* labels, size_t num_classes, double eps,
                      int metric, unsigned p, const void* queries, size_t q,
                      int32_t* out) {
  if (eps <= 0.0) return 4;
  size_t* hist = (size_t*)malloc(sizeof(size_t) * (num_classes + 1));
  for (size_t i = 0; i < q; ++i) {
    size_t nc = 0;
    int64_t* nb = radius_of(data, dtype, n, d, row_ptr(queries, dtype, d, i), eps, metric, p, &nc);
    for (size_t c = 0; c <= num_classes; ++c) hist[c] = 0;
    for (size_t j = 0; j < nc; ++j) {
      const int32_t yi = labels[nb[j]];
      if (yi < 0)
        hist[num_classes] += 1;
      else
        hist[yi] += 1;
    }
    size_t which = 0, m = hist[0];
    for (size_t c = 1; c <= num_classes; ++c)
      if (hist[c] > m) {
        m = hist[c];
        which = c;
      }
    out[i] = which != num_classes ? (int32_t)(which + 1) : 0;
    free(nb);
  }
  free(hist);
  return 0;
}
