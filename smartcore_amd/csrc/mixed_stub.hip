// Placeholder for the mixed-precision forms until their kernels land.
#include "common.hpp"
#include "kmeans.hpp"

namespace sc {

int knn_mixed_device(const sc_index*, const void*, size_t, size_t, int64_t*, double*, uint32_t*,
                     hipStream_t, sc_stats*) {
  return fail(SC_ERR_UNSUPPORTED, "SC_MIXED is not built in this library");
}

int kmeans_run_mixed(const sc_index*, int, size_t, double*, uint32_t*, uint64_t*, double*,
                     uint32_t*, hipStream_t) {
  return fail(SC_ERR_UNSUPPORTED, "SC_MIXED is not built in this library");
}

}  // namespace sc
