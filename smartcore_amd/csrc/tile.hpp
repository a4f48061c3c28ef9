// Tiled distance block shared by the exact kNN scan and the radius search.
//
// A 256-thread workgroup computes the 32 x 128 block of distance sums
// between queries q0..q0+31 and rows t0..t0+127.  Dimension chunks of 16
// are staged through LDS (transposed, one element of padding per LDS row to
// keep the transposed writes off a single bank); each thread folds a 4x4
// register micro-tile, every accumulator in the reference's element order.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "metric.hpp"

namespace sc {

constexpr int kTileQ = 32;
constexpr int kTileR = 128;
constexpr int kTileD = 16;
constexpr int kTileThreads = 256;

template <typename T>
struct TileSmem {
  T Qs[kTileD][kTileQ + 1];
  T Rs[kTileD][kTileR + 1];
};

// acc[i][j] = sum for query q0 + tq*4 + i and row t0 + tr + 32*j, where
// tq = tid >> 5 and tr = tid & 31.  Rows at or past r_end and queries at or
// past q contribute zeros (their sums are ignored by the callers).
template <typename T, int M>
__device__ __forceinline__ void tile_sums(TileSmem<T>& sm, double (&acc)[4][4],
                                          const T* __restrict__ refs, int64_t t0, int64_t r_end,
                                          const T* __restrict__ queries, int64_t q0, int64_t q,
                                          int ld, int d, double p) {
  const int tid = threadIdx.x;
  const int tq = tid >> 5;
  const int tr = tid & 31;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  // columns d .. ld are the row padding (zeros): their terms are exact
  // zeros added to non-negative sums, so stopping at d changes no bit
  for (int c0 = 0; c0 < d; c0 += kTileD) {
    const int kn = min(kTileD, d - c0);
    __syncthreads();
    for (int e = tid; e < kTileR * kTileD; e += kTileThreads) {
      const int r = e / kTileD, c = e % kTileD;
      const int64_t row = t0 + r;
      sm.Rs[c][r] = row < r_end ? refs[row * ld + c0 + c] : T(0);
    }
    for (int e = tid; e < kTileQ * kTileD; e += kTileThreads) {
      const int r = e / kTileD, c = e % kTileD;
      const int64_t row = q0 + r;
      sm.Qs[c][r] = row < q ? queries[row * ld + c0 + c] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kn; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sm.Qs[kk][tq * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sm.Rs[kk][tr + 32 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc[i][j] + Term<T, M>::f(a[i], b[j], p);
    }
  }
}

}  // namespace sc
