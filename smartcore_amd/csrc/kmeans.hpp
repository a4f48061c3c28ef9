// k-means device state and launch helpers shared by the FP64 and MFMA forms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace sc {

// Rows per partial-sum block.  Fixes the reduction order of the centroid
// sums and of the distortion: sums over each block's rows in row order, then
// over blocks in block order.
constexpr int kBlockRows = 2048;

// Device-side Lloyd loop state (kmeans.rs:273, :305-309).
struct KmState {
  double distortion;  // best distortion so far (f64::MAX at start)
  double dist;        // distortion of the last step
  uint32_t iters;     // steps executed
  uint32_t done;      // stop rule fired
  uint64_t fallback;  // rows reassigned in FP64 by the certified MFMA form
};

struct KmWork {
  DevBuf b_labels, b_mind, b_psum, b_pcount, b_pdist, b_state;
  uint32_t* labels = nullptr;
  double* mind = nullptr;
  double* psum = nullptr;
  uint32_t* pcount = nullptr;
  double* pdist = nullptr;
  KmState* state = nullptr;
  int alloc(const sc_index* ix, int k, hipStream_t s);
};

int kmeans_assign_launch(const sc_index* ix, const void* rows, int64_t n, const double* d_cent,
                         int k, uint32_t* d_labels, double* d_mind, const KmState* st,
                         hipStream_t s);
int kmeans_update_launch(const sc_index* ix, KmWork& w, int k, KmState* st, double* d_sums,
                         uint64_t* d_counts, double* d_cent, bool update, hipStream_t s);
int kmeans_state_init(KmState* st, hipStream_t s);
int kmeans_control_launch(KmState* st, hipStream_t s);

// The certified BF16x3-MFMA Lloyd loop (same contract as sc_kmeans_run).
int kmeans_run_mixed(const sc_index* ix, int k, size_t max_iter, double* centroids,
                     uint32_t* labels, uint64_t* size, double* distortion, uint32_t* iters,
                     hipStream_t s);

}  // namespace sc
