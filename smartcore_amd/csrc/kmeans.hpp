// k-means device state and launch helpers shared by the FP64 and MFMA forms
// and by the batched form.
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
// Rows a workgroup stages through LDS at a time in the FP64 form's partial
// kernel (4 waves x 32 rows).
constexpr int kChunkRows = 128;
// Largest k * d whose f64 accumulators fit the LDS partial kernels.
constexpr int kMaxLdsAcc = 8192;
// Elements of one cluster's sum vector handled by one reduce workgroup.
constexpr int kReduceChunk = 256;

// ||x - mu||^2 of one row over its ld columns in f64, the columns folded in
// order.  f32 rows with ld a multiple of 16 (every padded f32 row: its start
// is 64-byte aligned) are read 16 values per step; the fold is the same.
__device__ inline double row_shift_x2(const double* __restrict__ row,
                                      const double* __restrict__ mu, int ld) {
  double t = 0.0;
  for (int c = 0; c < ld; ++c) {
    const double v = row[c] - mu[c];
    t = t + v * v;
  }
  return t;
}
__device__ inline double row_shift_x2(const float* __restrict__ row,
                                      const double* __restrict__ mu, int ld) {
  double t = 0.0;
  if (ld % 16 != 0) {
    for (int c = 0; c < ld; ++c) {
      const double v = static_cast<double>(row[c]) - mu[c];
      t = t + v * v;
    }
    return t;
  }
  for (int c0 = 0; c0 < ld; c0 += 16) {
    float4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = *reinterpret_cast<const float4*>(row + c0 + 4 * u);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float e[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const double v = static_cast<double>(e[w]) - mu[c0 + 4 * u + w];
        t = t + v * v;
      }
    }
  }
  return t;
}

// Device-side Lloyd loop state (kmeans.rs:273, :305-309).
struct KmState {
  double distortion;  // best distortion so far (f64::MAX at start)
  double dist;        // distortion of the last step
  double x2total;     // sum over rows of ||x - mu||^2 (fixed order)
  uint32_t iters;     // steps executed
  uint32_t done;      // stop rule fired
  unsigned long long fallback;  // rows reassigned in FP64 by the certified MFMA form
};

// Everything the kernels of one k-means problem touch.  The batched form
// keeps an array of these in device memory; the single-problem form passes
// one by value.
struct KmProb {
  int64_t n, nb, ng;
  int d, ld, k, kpad;
  const void* rows;     // n x ld, f32 (or f64 in the FP64 form)
  double* cent;         // k x ld, the centroids of the current step
  uint32_t* labels;     // n
  double* psum;         // nb x k x d
  uint32_t* pcount;     // nb x k
  double* gsum;         // ng x k x d
  uint64_t* gcount;     // ng x k
  double* sums;         // k x d
  uint64_t* counts;     // k
  double* term;         // k distortion terms of the last step
  double* x2block;      // nb
  double* mu;           // ld: the shift of the distortion sums (the first row,
                        // non-finite values replaced by 0)
  KmState* st;
  uint32_t* arrive;     // arrival counter of the final reduce kernel
  // MFMA form only (null otherwise): the terms of the current centroids
  uint16_t* bhi;        // kpad x ld, BF16 RN(v), v = RN_f32(-2 c)
  uint16_t* bmd;        // kpad x ld, BF16 RN(v - bhi)
  uint16_t* blo;        // kpad x ld, BF16 RN(v - bhi - bmd)
  float* c2f;           // kpad, RN(||c||^2); +inf for padding rows
  double* c2d;          // k, ||c||^2 in f64
  float* cmax;          // RU(max ||c||), RU(max ||c||^2)
};

// Work items of the reduce kernels over a set of problems.
//   group items: (problem, group * slices + slice), a slice being 256
//                consecutive elements of the k x d sums
//   final items: (problem, cluster)
struct KmItems {
  DevBuf b_group, b_final;
  uint2* group = nullptr;
  uint2* fin = nullptr;
  int64_t ngroup = 0, nfinal = 0;
  // problems (the first nprob of the array); with more than one the final
  // kernel leaves the distortion and stop rule to a per-problem close kernel
  int64_t nprob = 0;
  // one problem whose group sums fit the final kernel's LDS: that kernel
  // forms them, and the group launch is left out
  bool fused = false;
  int build(const KmProb* probs, size_t count, hipStream_t s);
};

// Buffers of one problem: p points into them; dprob is p in device memory.
struct KmWork {
  DevBuf b_labels, b_psum, b_pcount, b_gsum, b_gcount, b_sums, b_counts, b_state, b_x2, b_tab,
      b_prob;
  KmProb p{};
  KmProb* dprob = nullptr;
  KmItems items;
  int alloc(const sc_index* ix, int k, const double* d_cent, bool mixed, hipStream_t s);
};

int kmeans_state_init(KmState* st, hipStream_t s);
// The shift point mu (the first row, widened) and the sum over the index's
// rows of ||x - mu||^2 in f64 into st->x2total.
int kmeans_x2_launch(const sc_index* ix, KmWork& w, hipStream_t s);
int kmeans_assign_launch(const sc_index* ix, const void* rows, int64_t n, const double* d_cent,
                         int k, uint32_t* d_labels, const KmState* st, hipStream_t s);
// Block partial sums of the rows by label (FP64 form).
int kmeans_partial_launch(const sc_index* ix, KmWork& w, int k, const KmState* st, hipStream_t s);
// Group and final sums, counts, distortion from the aggregates, shifted by
// mu so that data far from the origin lose no precision to cancellation:
//   D = sum_i ||x_i - mu||^2 - 2 sum_c (c - mu).(S_c - n_c mu)
//       + sum_c n_c ||c - mu||^2,
// and (update) centroid_i = sums_i / size_i for size_i > 0, (control) the
// stop rule, and the MFMA terms of the new centroids when the problem has
// them.  `probs` is a device array of `count` problems.
int kmeans_reduce_launch(const KmProb* d_probs, const KmItems& items, bool update, bool control,
                         hipStream_t s);
// MFMA terms of the current centroids (before the first step).
int kmeans_tables_launch(const KmProb* d_probs, size_t count, hipStream_t s);

// One certified MFMA assignment of every row of a single problem, then the
// block partial sums from the labels (same labels and partials as assign +
// partial of the FP64 form).
int kmeans_mixed_step_launch(const sc_index* ix, KmWork& w, hipStream_t s);
bool kmeans_mixed_supported(const sc_index* ix, int k);

}  // namespace sc
