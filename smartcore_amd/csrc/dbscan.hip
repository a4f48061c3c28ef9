// DBSCAN::fit and DBSCAN::predict (cluster/dbscan.rs:274-397) over a device
// index.
//
// The distance work -- find_radius (linear_search.rs:89-108) of every row,
// O(n^2 d) -- runs as one batched device radius search (each row's
// neighbours in index order, which is exactly the list find_radius returns).
//
// The reference's expansion loop (outer loop over the rows, `neighbors`
// used as a stack) gives these labels, which are computed here without the
// loop's order:
//   - core rows (at least min_samples neighbours, the row itself included)
//     joined by neighbour edges form the clusters, and the outer loop starts
//     each at its smallest core row, so cluster k is the component whose
//     smallest core row is the k-th such row in index order;
//   - every other row takes the first cluster whose expansion reaches it,
//     i.e. the lowest-numbered cluster among its core neighbours (clusters
//     are expanded one at a time, in order), or -1 when it has none.
// The components come from a union-find on the device over the core rows'
// edges (links always point a root at a smaller root, so a component's root
// is its smallest row); the numbering and the non-core rows (whose lists are
// shorter than min_samples) are done on the host.  predict searches its
// queries the same way and takes the first most frequent label (which_max,
// tree/decision_tree_classifier.rs:490-502).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

using namespace sc;

namespace {

// The root of v's tree, halving the path on the way.  parent[x] <= x holds
// for every value ever stored, so reads that are stale still point at an
// ancestor, and the walk strictly decreases until it meets a root.
__device__ uint32_t dbscan_rep(uint32_t* parent, uint32_t v) {
  uint32_t cur = parent[v];
  if (cur != v) {
    uint32_t prev = v, next;
    while (cur > (next = parent[cur])) {
      parent[prev] = next;
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

// Joins the trees of u and v: the larger root is linked to the smaller by a
// compare-and-swap; a failed swap returns a smaller value to continue from,
// so the loop ends.
__device__ void dbscan_link(uint32_t* parent, uint32_t u, uint32_t v) {
  uint32_t ru = dbscan_rep(parent, u), rv = dbscan_rep(parent, v);
  while (ru != rv) {
    if (ru < rv) {
      const uint32_t old = atomicCAS(&parent[rv], rv, ru);
      if (old == rv) break;
      rv = old;
    } else {
      const uint32_t old = atomicCAS(&parent[ru], ru, rv);
      if (old == ru) break;
      ru = old;
    }
  }
}

__global__ void dbscan_init_kernel(const uint64_t* __restrict__ off, int64_t n,
                                   uint64_t min_samples, uint8_t* __restrict__ core,
                                   uint32_t* __restrict__ parent) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    core[i] = off[i + 1] - off[i] >= min_samples ? 1 : 0;
    parent[i] = static_cast<uint32_t>(i);
  }
}

// One wavefront per row: a core row is joined with its core neighbours of
// smaller index (the relation is symmetric, so each edge is taken once).
__global__ void dbscan_union_kernel(const uint64_t* __restrict__ off,
                                    const int64_t* __restrict__ idx, int64_t n,
                                    const uint8_t* __restrict__ core, uint32_t* parent) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; i < n;
       i += waves) {
    if (!core[i]) continue;
    const uint64_t e1 = off[i + 1];
    for (uint64_t e = off[i] + lane; e < e1; e += 64) {
      const int64_t j = idx[e];
      if (j < i && core[j]) dbscan_link(parent, static_cast<uint32_t>(i), static_cast<uint32_t>(j));
    }
  }
}

// The component root of every core row (0xffffffff for the others).
__global__ void dbscan_root_kernel(int64_t n, const uint8_t* __restrict__ core, uint32_t* parent,
                                   uint32_t* __restrict__ root) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    root[i] = core[i] ? dbscan_rep(parent, static_cast<uint32_t>(i)) : 0xffffffffu;
}

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
  const size_t total = nb.idx.size();
  std::vector<uint32_t> root(n);
  {
    DevBuf doff, didx, dcore, dpar, droot;
    SC_HIP(doff.alloc(sizeof(uint64_t) * (n + 1), s));
    SC_HIP(didx.alloc(sizeof(int64_t) * std::max<size_t>(total, 1), s));
    SC_HIP(dcore.alloc(n, s));
    SC_HIP(dpar.alloc(sizeof(uint32_t) * n, s));
    SC_HIP(droot.alloc(sizeof(uint32_t) * n, s));
    if (int rc = upload_staged(doff.p, nb.offsets.data(), sizeof(uint64_t) * (n + 1), s)) return rc;
    if (int rc = upload_staged(didx.p, nb.idx.data(), sizeof(int64_t) * total, s)) return rc;
    const int64_t nn = static_cast<int64_t>(n);
    const unsigned g = static_cast<unsigned>(std::min<int64_t>(4096, (nn + 255) / 256));
    hipLaunchKernelGGL(dbscan_init_kernel, dim3(g), dim3(256), 0, s, doff.as<uint64_t>(), nn,
                       static_cast<uint64_t>(min_samples), dcore.as<uint8_t>(), dpar.as<uint32_t>());
    const unsigned gu = static_cast<unsigned>(std::min<int64_t>(8192, (nn + 3) / 4));
    hipLaunchKernelGGL(dbscan_union_kernel, dim3(gu), dim3(256), 0, s, doff.as<uint64_t>(),
                       didx.as<int64_t>(), nn, dcore.as<uint8_t>(), dpar.as<uint32_t>());
    hipLaunchKernelGGL(dbscan_root_kernel, dim3(g), dim3(256), 0, s, nn, dcore.as<uint8_t>(),
                       dpar.as<uint32_t>(), droot.as<uint32_t>());
    SC_HIP(hipGetLastError());
    if (int rc = download_staged(root.data(), droot.p, sizeof(uint32_t) * n, s)) return rc;
  }
  auto count = [&](size_t i) { return static_cast<size_t>(nb.offsets[i + 1] - nb.offsets[i]); };
  // clusters numbered by their smallest core row, in index order
  // (dbscan.rs:286-346: the outer loop starts a cluster there)
  std::vector<int32_t> cid(n, -1);
  int32_t k = 0;
  for (size_t i = 0; i < n; ++i)
    if (count(i) >= min_samples && root[i] == i) cid[i] = k++;
  std::vector<int32_t> y(n, -1);
  for (size_t i = 0; i < n; ++i) {
    if (count(i) >= min_samples) {
      y[i] = cid[root[i]];
      continue;
    }
    // a non-core row: the lowest cluster among its core neighbours
    int32_t best = -1;
    for (uint64_t e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
      const size_t j = static_cast<size_t>(nb.idx[e]);
      if (count(j) >= min_samples) {
        const int32_t c = cid[root[j]];
        if (best < 0 || c < best) best = c;
      }
    }
    y[i] = best;
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
