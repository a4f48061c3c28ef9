// k-means device state and launch helpers shared by the FP64 and MFMA forms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace sc {

// Rows per partial-sum block.  Fixes the reduction order of the centroid
// sums: each cluster's rows are added in row order inside a block (starting
// from 0), block sums are added in groups of kGroupBlocks consecutive blocks,
// and group sums in group order.  Nothing depends on the launch shape.
constexpr int kBlockRows = 2048;
constexpr int kGroupBlocks = 16;
// Rows a workgroup stages through LDS at a time in the partial and fused
// kernels (4 waves x 32 rows).
constexpr int kChunkRows = 128;
// Largest k * d whose f64 accumulators fit the LDS partial kernels.
constexpr int kMaxLdsAcc = 8192;

// Device-side Lloyd loop state (kmeans.rs:273, :305-309).
struct KmState {
  double distortion;  // best distortion so far (f64::MAX at start)
  double dist;        // distortion of the last step
  double x2total;     // sum over rows of ||x||^2 (fixed order)
  uint32_t iters;     // steps executed
  uint32_t done;      // stop rule fired
  unsigned long long fallback;  // rows reassigned in FP64 by the certified MFMA form
};

struct KmWork {
  DevBuf b_labels, b_psum, b_pcount, b_gsum, b_gcount, b_sums, b_counts, b_state, b_x2;
  uint32_t* labels = nullptr;
  double* psum = nullptr;      // nblocks x k x d
  uint32_t* pcount = nullptr;  // nblocks x k
  double* gsum = nullptr;      // ngroups x k x d
  uint64_t* gcount = nullptr;  // ngroups x k
  double* sums = nullptr;      // k x d
  uint64_t* counts = nullptr;  // k
  double* x2block = nullptr;   // nblocks
  KmState* state = nullptr;
  int64_t nblocks = 0, ngroups = 0;
  int alloc(const sc_index* ix, int k, hipStream_t s);
};

int kmeans_state_init(KmState* st, hipStream_t s);
// Sum over the index's rows of ||x||^2 in f64 into st->x2total.
int kmeans_x2_launch(const sc_index* ix, KmWork& w, hipStream_t s);
int kmeans_assign_launch(const sc_index* ix, const void* rows, int64_t n, const double* d_cent,
                         int k, uint32_t* d_labels, const KmState* st, hipStream_t s);
// Block partial sums of the rows by label (FP64 form).
int kmeans_partial_launch(const sc_index* ix, KmWork& w, int k, const KmState* st, hipStream_t s);
// Group and final sums, counts, distortion from the aggregates, and (when
// `update`) centroid_i = sums_i / size_i for size_i > 0.
int kmeans_finish_launch(const sc_index* ix, KmWork& w, int k, KmState* st, double* d_cent,
                         bool update, hipStream_t s);
int kmeans_control_launch(KmState* st, hipStream_t s);

// One certified MFMA assignment + fused partial sums (same outputs as
// assign + partial of the FP64 form).
int kmeans_mixed_step_launch(const sc_index* ix, KmWork& w, int k, const double* d_cent,
                             KmState* st, hipStream_t s);
bool kmeans_mixed_supported(const sc_index* ix, int k);

}  // namespace sc
