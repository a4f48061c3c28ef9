// DBSCAN::fit and DBSCAN::predict (cluster/dbscan.rs:274-397) over a device
// index.
//
// The distance work -- find_radius (linear_search.rs:89-108) of every row,
// O(n^2 d) -- runs as one batched device radius search (count and fill
// passes; each row's neighbours in index order, which is exactly the list
// find_radius returns).  The reference's expansion loop then runs on the
// host over those lists in the reference's order (outer loop over the rows,
// `neighbors` used as a stack), so the cluster labels and the number of
// clusters are the reference's.  predict searches its queries the same way
// and takes the first most frequent label (which_max,
// tree/decision_tree_classifier.rs:490-502).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

using namespace sc;

namespace {

int check_metric(sc_metric metric, uint16_t p) {
  if (metric != SC_EUCLIDIAN && metric != SC_MANHATTAN && metric != SC_MINKOWSKI)
    return fail(SC_ERR_UNSUPPORTED, "unknown metric");
  if (metric == SC_MINKOWSKI && p < 1) return fail(SC_ERR_PANIC, "p must be at least 1");
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_dbscan_fit(const sc_index* index, double eps, size_t min_samples, sc_metric metric,
                  uint16_t p, int32_t* labels, size_t* num_classes) {
  clear_error();
  if (!index || !labels || !num_classes) return fail(SC_PARAMETERS_ERROR, "null argument");
  *num_classes = 0;
  // dbscan.rs:278-284
  if (min_samples < 1) return fail(SC_FIT_FAILED, "Invalid minPts");
  if (eps <= 0.0) return fail(SC_FIT_FAILED, "Invalid radius: ");
  if (int rc = check_metric(metric, p)) return rc;
  const size_t n = index->n;
  if (n == 0) return SC_OK;
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  sc_radius_result nb;
  // the expansion reads the neighbour rows only, not their distances
  if (int rc = radius_search_device(index, index->rows, static_cast<int64_t>(n), metric, p, eps,
                                    &nb, s, false))
    return rc;
  auto count = [&](size_t i) { return static_cast<size_t>(nb.offsets[i + 1] - nb.offsets[i]); };
  auto at = [&](size_t i, size_t j) { return nb.idx[nb.offsets[i] + j]; };

  // dbscan.rs:286-346
  const int32_t queued = -2, outlier = -1, undefined = -3;
  std::vector<int32_t> y(n, undefined);
  int32_t k = 0;
  std::vector<int64_t> stack;
  for (size_t i = 0; i < n; ++i) {
    if (y[i] != undefined) continue;
    const size_t nc = count(i);
    if (nc < min_samples) {
      y[i] = outlier;
      continue;
    }
    y[i] = k;
    stack.clear();
    for (size_t j = 0; j < nc; ++j) {
      const int64_t r = at(i, j);
      if (y[r] == undefined) y[r] = queued;
      stack.push_back(r);
    }
    while (!stack.empty()) {
      const int64_t index_r = stack.back();
      stack.pop_back();
      if (y[index_r] == outlier) y[index_r] = k;
      if (y[index_r] == undefined || y[index_r] == queued) {
        y[index_r] = k;
        const size_t sc = count(static_cast<size_t>(index_r));
        if (sc >= min_samples) {
          for (size_t j = 0; j < sc; ++j) {
            const int64_t r = at(static_cast<size_t>(index_r), j);
            const int32_t label = y[r];
            if (label == undefined) y[r] = queued;
            if (label == undefined || label == outlier) stack.push_back(r);
          }
        }
      }
    }
    k += 1;
  }
  std::copy(y.begin(), y.end(), labels);
  *num_classes = static_cast<size_t>(k);
  return SC_OK;
}

int sc_dbscan_predict(const sc_index* index, const int32_t* labels, size_t num_classes,
                      double eps, sc_metric metric, uint16_t p, const void* queries, size_t q,
                      int32_t* out) {
  clear_error();
  if (!index || !labels || !out || (!queries && q > 0))
    return fail(SC_PARAMETERS_ERROR, "null argument");
  // find_radius's own check (linear_search.rs:90-95)
  if (!(eps > 0.0)) return fail(SC_FIND_FAILED, "radius should be > 0");
  if (int rc = check_metric(metric, p)) return rc;
  if (q == 0) return SC_OK;
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  sc_radius_result nb;
  {
    DevBuf dq;
    SC_HIP(dq.alloc(q * index->ld * elem_size(index->dtype), s));
    if (int rc = upload_padded(queries, index->dtype, q, index->d, index->ld, dq.p, s)) return rc;
    if (int rc = radius_search_device(index, dq.p, static_cast<int64_t>(q), metric, p, eps, &nb, s,
                                      false))
      return rc;
  }
  // dbscan.rs:368-393
  std::vector<size_t> hist(num_classes + 1);
  for (size_t i = 0; i < q; ++i) {
    std::fill(hist.begin(), hist.end(), 0);
    for (uint64_t e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
      const int32_t yi = labels[nb.idx[e]];
      if (yi < 0)
        hist[num_classes] += 1;
      else if (static_cast<size_t>(yi) < num_classes)
        hist[static_cast<size_t>(yi)] += 1;
      else
        return fail(SC_PARAMETERS_ERROR, "label out of range");
    }
    size_t which = 0, m = hist[0];
    for (size_t c = 1; c <= num_classes; ++c)
      if (hist[c] > m) {
        m = hist[c];
        which = c;
      }
    out[i] = which != num_classes ? static_cast<int32_t>(which + 1) : 0;
  }
  return SC_OK;
}

}  // extern "C"
