// Placeholder for the certified MFMA k-means form until its kernels land.
#include "common.hpp"
#include "kmeans.hpp"

namespace sc {

int kmeans_run_mixed(const sc_index*, int, size_t, double*, uint32_t*, uint64_t*, double*,
                     uint32_t*, hipStream_t) {
  return fail(SC_ERR_UNSUPPORTED, "SC_MIXED k-means is not built in this library");
}

}  // namespace sc
