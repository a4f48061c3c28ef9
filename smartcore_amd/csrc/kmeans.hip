// k-means Lloyd step (BBDTree::clustering, bbd_tree.rs:62-87) and the loop of
// KMeans::fit (kmeans.rs:294-310), FP64 form, plus the reductions the MFMA
// form shares.
//
//   assign    Tiled FP64 squared distances of every row (widened to f64 as
//             KMeans::predict does, kmeans.rs:338-342) to every centroid, in
//             the reference's element order; argmin with strict < so the
//             lowest index wins ties (kmeans.rs:343-346).  64 rows x 64
//             centroids per tile, 4x4 register micro-tiles.
//   partial   One workgroup per block of kBlockRows rows.  Rows are staged
//             through LDS 128 at a time; wave w owns the clusters c with
//             c % 4 == w and adds each owned row into an f64 LDS accumulator
//             (lane = dimension, ds_add_f64), in row order.  One owner per
//             cluster makes the sums deterministic.
//   finish    Block sums added in groups of 16 blocks, groups in order;
//             distortion from the aggregates
//               D = sum_i ||x_i||^2 - 2 sum_c c.S_c + sum_c n_c ||c||^2
//             (the cost BBDTree::clustering returns, bbd_tree.rs:162 and
//             :297-305, is the same quantity); centroid_i = sums_i / size_i
//             when size_i > 0 (kmeans.rs:297-303).
//   control   The stop rule `if distortion <= dist break` (kmeans.rs:305-309)
//             on the device; later kernels return at once, so a fit needs
//             one host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "common.hpp"
#include "kmeans.hpp"
#include "metric.hpp"
#include "smallrng.hpp"

namespace sc {
namespace {

constexpr int PT = 64;  // rows per assign tile
constexpr int CT = 64;  // centroids per assign tile
constexpr int DCK = 16;
constexpr int NT = 256;

__device__ __forceinline__ void lexmin(double& v, uint32_t& i, double v2, uint32_t i2) {
  if (v2 < v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void kmeans_assign_kernel(const T* __restrict__ rows, int64_t n,
                                                           int ld, const double* __restrict__ cent,
                                                           int k, uint32_t* __restrict__ labels,
                                                           const KmState* __restrict__ st) {
  if (st && st->done) return;
  __shared__ double Ps[DCK][PT + 1];
  __shared__ double Cs[DCK][CT + 1];
  const int tid = threadIdx.x;
  const int tp = tid >> 4;  // rows tp*4 .. tp*4+3
  const int tc = tid & 15;  // centroids tc + 16*j
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * PT;

  double bv[4];
  uint32_t bi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bv[i] = INFINITY;
    bi[i] = 0xffffffffu;
  }
  for (int ct0 = 0; ct0 < k; ct0 += CT) {
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int c0 = 0; c0 < ld; c0 += DCK) {
      __syncthreads();
      for (int e = tid; e < PT * DCK; e += NT) {
        const int r = e / DCK, c = e % DCK;
        const int64_t row = p0 + r;
        Ps[c][r] = row < n ? static_cast<double>(rows[row * ld + c0 + c]) : 0.0;
      }
      for (int e = tid; e < CT * DCK; e += NT) {
        const int r = e / DCK, c = e % DCK;
        const int cr = ct0 + r;
        Cs[c][r] = cr < k ? cent[static_cast<int64_t>(cr) * ld + c0 + c] : 0.0;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < DCK; ++kk) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Ps[kk][tp * 4 + i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Cs[kk][tc + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = acc[i][j] + Term<double, kEuclid>::f(a[i], b[j], 2.0);
      }
    }
    // best of this tile per row: own 4 centroids in index order, then the
    // 16 lanes sharing the row (lexicographic (value, index) minimum equals
    // the first strict minimum of a sequential scan).
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double v = INFINITY;
      uint32_t ix = 0xffffffffu;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = ct0 + tc + 16 * j;
        if (c < k) lexmin(v, ix, acc[i][j], static_cast<uint32_t>(c));
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 16);
        const uint32_t i2 = __shfl_xor(ix, off, 16);
        lexmin(v, ix, v2, i2);
      }
      if (v < bv[i]) {
        bv[i] = v;
        bi[i] = ix;
      }
    }
  }
  if (tc == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = p0 + tp * 4 + i;
      // kmeans.rs:334-347: min starts at f64::MAX; nothing below it keeps 0
      if (row < n) labels[row] = bv[i] < DBL_MAX ? bi[i] : 0u;
    }
  }
}

// Per block: sum of ||x||^2 over its rows (each row a sequential f64 fold,
// 8 consecutive rows per thread, then a fixed tree).
template <typename T>
__global__ __launch_bounds__(NT) void x2_block_kernel(const T* __restrict__ rows, int64_t n, int ld,
                                                      double* __restrict__ x2block) {
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kBlockRows;
  double s = 0.0;
  for (int j = 0; j < kBlockRows / NT; ++j) {
    const int64_t r = r0 + tid * (kBlockRows / NT) + j;
    if (r < n) {
      double t = 0.0;
      for (int c = 0; c < ld; ++c) {
        const double v = static_cast<double>(rows[r * ld + c]);
        t = t + v * v;
      }
      s = s + t;
    }
  }
  red[tid] = s;
  __syncthreads();
  for (int w = NT / 2; w >= 1; w >>= 1) {
    if (tid < w) red[tid] = red[tid] + red[tid + w];
    __syncthreads();
  }
  if (tid == 0) x2block[blockIdx.x] = red[0];
}

__global__ void x2_total_kernel(const double* __restrict__ x2block, int64_t nb, KmState* st) {
  double s = 0.0;
  for (int64_t b = 0; b < nb; ++b) s = s + x2block[b];
  st->x2total = s;
}

// Partial sums with one owner wave per cluster (see the file comment).
// Dynamic LDS: acc[k*d] f64 | stage[kChunkRows*ld] T | lab[kChunkRows] | cnt[k].
template <typename T>
__global__ __launch_bounds__(NT) void kmeans_partial_lds_kernel(
    const T* __restrict__ rows, int64_t n, int ld, int d, const uint32_t* __restrict__ labels,
    int k, double* __restrict__ psum, uint32_t* __restrict__ pcount,
    const KmState* __restrict__ st) {
  if (st && st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = reinterpret_cast<double*>(smem);
  T* stage = reinterpret_cast<T*>(acc + (static_cast<size_t>(k) * d + 1) / 2 * 2);
  uint32_t* lab = reinterpret_cast<uint32_t*>(stage + static_cast<size_t>(kChunkRows) * ld);
  uint32_t* cnt = lab + kChunkRows;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t b = blockIdx.x;
  const int64_t r0 = b * kBlockRows;
  const int rows_here = static_cast<int>(min<int64_t>(kBlockRows, n - r0));
  for (int e = tid; e < k * d; e += NT) acc[e] = 0.0;
  for (int e = tid; e < k; e += NT) cnt[e] = 0u;
  for (int c0 = 0; c0 < rows_here; c0 += kChunkRows) {
    const int nr = min(kChunkRows, rows_here - c0);
    __syncthreads();
    const T* src = rows + (r0 + c0) * ld;
    for (int e = tid; e < nr * ld; e += NT) stage[e] = src[e];
    for (int e = tid; e < kChunkRows; e += NT)
      lab[e] = e < nr ? labels[r0 + c0 + e] : 0xffffffffu;
    __syncthreads();
    for (int i0 = 0; i0 < nr; i0 += 64) {
      const uint32_t li = lab[i0 + lane];
      const bool mine = i0 + lane < nr && (li & 3u) == static_cast<uint32_t>(wave);
      uint64_t m = __ballot(mine);
      while (m) {
        const int i = i0 + __builtin_ctzll(m);
        m &= m - 1;
        const uint32_t c = lab[i];
        for (int dc = 0; dc < d; dc += 64) {
          const int dim = dc + lane;
          if (dim < d)
            atomicAdd(&acc[static_cast<size_t>(c) * d + dim],
                      static_cast<double>(stage[static_cast<size_t>(i) * ld + dim]));
        }
        if (lane == 0) cnt[c] += 1u;
      }
    }
  }
  __syncthreads();
  double* out = psum + b * static_cast<int64_t>(k) * d;
  for (int e = tid; e < k * d; e += NT) out[e] = acc[e];
  for (int e = tid; e < k; e += NT) pcount[b * k + e] = cnt[e];
}

// Fallback when the accumulators do not fit LDS: one wave per cluster,
// ballots over the labels, rows added in row order.
template <typename T>
__global__ __launch_bounds__(NT) void kmeans_partial_ballot_kernel(
    const T* __restrict__ rows, int64_t n, int ld, int d, const uint32_t* __restrict__ labels,
    int k, double* __restrict__ psum, uint32_t* __restrict__ pcount,
    const KmState* __restrict__ st) {
  if (st && st->done) return;
  __shared__ uint32_t lab[kBlockRows];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t b = blockIdx.x;
  const int64_t r0 = b * kBlockRows;
  const int rows_here = static_cast<int>(min<int64_t>(kBlockRows, n - r0));
  for (int i = tid; i < kBlockRows; i += NT) lab[i] = i < rows_here ? labels[r0 + i] : 0xffffffffu;
  __syncthreads();
  double* my_psum = psum + b * static_cast<int64_t>(k) * d;
  for (int c = wave; c < k; c += NT / 64) {
    uint32_t cnt = 0;
    for (int dc0 = 0; dc0 < d; dc0 += 64) {
      const int dim = dc0 + lane;
      const bool live = dim < d;
      double acc = 0.0;
      for (int i0 = 0; i0 < rows_here; i0 += 64) {
        const int i = i0 + lane;
        const bool mine = i < rows_here && lab[i] == static_cast<uint32_t>(c);
        uint64_t m = __ballot(mine);
        if (dc0 == 0) cnt += static_cast<uint32_t>(__popcll(m));
        while (m) {
          const int r = __builtin_ctzll(m);
          m &= m - 1;
          if (live) acc = acc + static_cast<double>(rows[(r0 + i0 + r) * ld + dim]);
        }
      }
      if (live) my_psum[static_cast<int64_t>(c) * d + dim] = acc;
    }
    if (lane == 0) pcount[b * k + c] = cnt;
  }
}

// gsum[g][c][j] = sum of psum over blocks 16g .. 16g+15, in block order.
__global__ void kmeans_group_kernel(int64_t nb, int64_t ng, int k, int d,
                                    const double* __restrict__ psum,
                                    const uint32_t* __restrict__ pcount, double* __restrict__ gsum,
                                    uint64_t* __restrict__ gcount, const KmState* __restrict__ st) {
  if (st && st->done) return;
  const int64_t kd = static_cast<int64_t>(k) * d;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ng * kd) return;
  const int64_t g = t / kd, e = t % kd;
  const int64_t b0 = g * kGroupBlocks, b1 = min<int64_t>(nb, b0 + kGroupBlocks);
  double s = 0.0;
  for (int64_t b = b0; b < b1; ++b) s = s + psum[b * kd + e];
  gsum[g * kd + e] = s;
  if (e % d == 0) {
    const int64_t c = e / d;
    uint64_t cnt = 0;
    for (int64_t b = b0; b < b1; ++b) cnt += pcount[b * k + c];
    gcount[g * k + c] = cnt;
  }
}

__global__ void kmeans_sum_kernel(int64_t ng, int k, int d, const double* __restrict__ gsum,
                                  const uint64_t* __restrict__ gcount, double* __restrict__ sums,
                                  uint64_t* __restrict__ counts, const KmState* __restrict__ st) {
  if (st && st->done) return;
  const int64_t kd = static_cast<int64_t>(k) * d;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= kd) return;
  double s = 0.0;
  for (int64_t g = 0; g < ng; ++g) s = s + gsum[g * kd + e];
  sums[e] = s;
  if (e % d == 0) {
    const int64_t c = e / d;
    uint64_t cnt = 0;
    for (int64_t g = 0; g < ng; ++g) cnt += gcount[g * k + c];
    counts[c] = cnt;
  }
}

// One workgroup: distortion from the aggregates with the centroids of this
// step, then (update) the new centroids.
__global__ __launch_bounds__(NT) void kmeans_distortion_update_kernel(
    int k, int d, int ld, const double* __restrict__ sums, const uint64_t* __restrict__ counts,
    double* __restrict__ cent, double* __restrict__ term, KmState* __restrict__ st, int update) {
  if (st->done) return;
  const int tid = threadIdx.x;
  for (int c = tid; c < k; c += NT) {
    const double* cc = cent + static_cast<int64_t>(c) * ld;
    const double* sc_ = sums + static_cast<int64_t>(c) * d;
    double dot = 0.0, c2 = 0.0;
    for (int j = 0; j < d; ++j) {
      dot = dot + cc[j] * sc_[j];
      c2 = c2 + cc[j] * cc[j];
    }
    term[c] = counts[c] > 0 ? static_cast<double>(counts[c]) * c2 - 2.0 * dot : 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    double s = st->x2total;
    for (int c = 0; c < k; ++c) s = s + term[c];
    st->dist = s;
  }
  __syncthreads();
  if (update) {
    for (int e = tid; e < k * d; e += NT) {
      const int c = e / d, j = e % d;
      if (counts[c] > 0)
        cent[static_cast<int64_t>(c) * ld + j] = sums[e] / static_cast<double>(counts[c]);
    }
  }
}

__global__ void kmeans_control_kernel(KmState* st) {
  if (st->done) return;
  st->iters += 1;
  if (st->distortion <= st->dist) {
    st->done = 1;
  } else {
    st->distortion = st->dist;
  }
}

__global__ void kmeans_state_init_kernel(KmState* st) {
  st->distortion = DBL_MAX;
  st->dist = 0.0;
  st->x2total = 0.0;
  st->iters = 0;
  st->done = 0;
  st->fallback = 0;
}

// One k-means++ round (kmeans.rs:367-379 and the final pass :399-410):
// d_i = min(d_i, squared_distance(row_i, centroid)) in the rows' own type
// (the reference calls Euclidian::squared_distance on TX rows here), and the
// label of a row that improves becomes `label`.
template <typename T>
__global__ void kmeanspp_round_kernel(const T* __restrict__ rows, int64_t n, int ld,
                                      const T* __restrict__ centroid, double* __restrict__ dmin,
                                      uint32_t* __restrict__ y, uint32_t label) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* r = rows + i * ld;
  double s = 0.0;
  for (int j = 0; j < ld; ++j) s = s + Term<T, kEuclid>::f(r[j], centroid[j], 2.0);
  if (s < dmin[i]) {
    dmin[i] = s;
    y[i] = label;
  }
}

__global__ void fill_kernel(double* p, int64_t n, double v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

size_t partial_lds_bytes(const sc_index* ix, int k) {
  return sizeof(double) * ((static_cast<size_t>(k) * ix->d + 1) / 2 * 2) +
         elem_size(ix->dtype) * static_cast<size_t>(kChunkRows) * ix->ld +
         sizeof(uint32_t) * (kChunkRows + static_cast<size_t>(k)) + 16;
}

}  // namespace

int kmeans_assign_launch(const sc_index* ix, const void* rows, int64_t n, const double* d_cent,
                         int k, uint32_t* d_labels, const KmState* st, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((n + PT - 1) / PT);
  if (grid == 0) return SC_OK;
  ProfScope ps("kmeans_assign", s);
  if (ix->dtype == SC_F64)
    hipLaunchKernelGGL(kmeans_assign_kernel<double>, dim3(grid), dim3(NT), 0, s,
                       static_cast<const double*>(rows), n, static_cast<int>(ix->ld), d_cent, k,
                       d_labels, st);
  else
    hipLaunchKernelGGL(kmeans_assign_kernel<float>, dim3(grid), dim3(NT), 0, s,
                       static_cast<const float*>(rows), n, static_cast<int>(ix->ld), d_cent, k,
                       d_labels, st);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int kmeans_x2_launch(const sc_index* ix, KmWork& w, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  if (w.nblocks > 0) {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(x2_block_kernel<double>, dim3(static_cast<unsigned>(w.nblocks)), dim3(NT),
                         0, s, static_cast<const double*>(ix->rows), n, static_cast<int>(ix->ld),
                         w.x2block);
    else
      hipLaunchKernelGGL(x2_block_kernel<float>, dim3(static_cast<unsigned>(w.nblocks)), dim3(NT),
                         0, s, static_cast<const float*>(ix->rows), n, static_cast<int>(ix->ld),
                         w.x2block);
    SC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(x2_total_kernel, dim3(1), dim3(1), 0, s, w.x2block, w.nblocks, w.state);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int kmeans_partial_launch(const sc_index* ix, KmWork& w, int k, const KmState* st, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  const int d = static_cast<int>(ix->d);
  const int ld = static_cast<int>(ix->ld);
  if (w.nblocks == 0) return SC_OK;
  ProfScope ps("kmeans_partial", s);
  const size_t lds = partial_lds_bytes(ix, k);
  const dim3 grid(static_cast<unsigned>(w.nblocks));
  if (static_cast<int64_t>(k) * d <= kMaxLdsAcc && lds <= 96 * 1024) {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(kmeans_partial_lds_kernel<double>, grid, dim3(NT), lds, s,
                         static_cast<const double*>(ix->rows), n, ld, d, w.labels, k, w.psum,
                         w.pcount, st);
    else
      hipLaunchKernelGGL(kmeans_partial_lds_kernel<float>, grid, dim3(NT), lds, s,
                         static_cast<const float*>(ix->rows), n, ld, d, w.labels, k, w.psum,
                         w.pcount, st);
  } else {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(kmeans_partial_ballot_kernel<double>, grid, dim3(NT), 0, s,
                         static_cast<const double*>(ix->rows), n, ld, d, w.labels, k, w.psum,
                         w.pcount, st);
    else
      hipLaunchKernelGGL(kmeans_partial_ballot_kernel<float>, grid, dim3(NT), 0, s,
                         static_cast<const float*>(ix->rows), n, ld, d, w.labels, k, w.psum,
                         w.pcount, st);
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int kmeans_finish_launch(const sc_index* ix, KmWork& w, int k, KmState* st, double* d_cent,
                         bool update, hipStream_t s) {
  const int d = static_cast<int>(ix->d);
  const int64_t kd = std::max<int64_t>(1, static_cast<int64_t>(k) * d);
  ProfScope ps("kmeans_reduce", s);
  if (w.ngroups > 0) {
    const int64_t t = w.ngroups * kd;
    hipLaunchKernelGGL(kmeans_group_kernel, dim3(static_cast<unsigned>((t + 255) / 256)), dim3(256),
                       0, s, w.nblocks, w.ngroups, k, d, w.psum, w.pcount, w.gsum, w.gcount, st);
    SC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(kmeans_sum_kernel, dim3(static_cast<unsigned>((kd + 255) / 256)), dim3(256), 0,
                     s, w.ngroups, k, d, w.gsum, w.gcount, w.sums, w.counts, st);
  SC_HIP(hipGetLastError());
  hipLaunchKernelGGL(kmeans_distortion_update_kernel, dim3(1), dim3(NT), 0, s, k, d,
                     static_cast<int>(ix->ld), w.sums, w.counts, d_cent,
                     w.x2block + std::max<int64_t>(w.nblocks, 1), st, update ? 1 : 0);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int KmWork::alloc(const sc_index* ix, int k, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  const int64_t d = std::max<int64_t>(static_cast<int64_t>(ix->d), 1);
  nblocks = (n + kBlockRows - 1) / kBlockRows;
  ngroups = (nblocks + kGroupBlocks - 1) / kGroupBlocks;
  const int64_t nb1 = std::max<int64_t>(nblocks, 1), ng1 = std::max<int64_t>(ngroups, 1);
  SC_HIP(b_labels.alloc(sizeof(uint32_t) * std::max<int64_t>(n, 1), s));
  SC_HIP(b_psum.alloc(sizeof(double) * nb1 * k * d, s));
  SC_HIP(b_pcount.alloc(sizeof(uint32_t) * nb1 * k, s));
  SC_HIP(b_gsum.alloc(sizeof(double) * ng1 * k * d, s));
  SC_HIP(b_gcount.alloc(sizeof(uint64_t) * ng1 * k, s));
  SC_HIP(b_sums.alloc(sizeof(double) * k * d, s));
  SC_HIP(b_counts.alloc(sizeof(uint64_t) * k, s));
  // x2 per block, then k distortion terms
  SC_HIP(b_x2.alloc(sizeof(double) * (nb1 + k), s));
  SC_HIP(b_state.alloc(sizeof(KmState), s));
  labels = b_labels.as<uint32_t>();
  psum = b_psum.as<double>();
  pcount = b_pcount.as<uint32_t>();
  gsum = b_gsum.as<double>();
  gcount = b_gcount.as<uint64_t>();
  sums = b_sums.as<double>();
  counts = b_counts.as<uint64_t>();
  x2block = b_x2.as<double>();
  state = b_state.as<KmState>();
  return SC_OK;
}

int kmeans_state_init(KmState* st, hipStream_t s) {
  hipLaunchKernelGGL(kmeans_state_init_kernel, dim3(1), dim3(1), 0, s, st);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int kmeans_control_launch(KmState* st, hipStream_t s) {
  hipLaunchKernelGGL(kmeans_control_kernel, dim3(1), dim3(1), 0, s, st);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc

using namespace sc;

namespace {

int upload_centroids(const double* host, int k, size_t d, size_t ld, double* dev, hipStream_t s) {
  SC_HIP(hipMemsetAsync(dev, 0, sizeof(double) * k * ld, s));
  if (d > 0)
    SC_HIP(hipMemcpy2DAsync(dev, ld * sizeof(double), host, d * sizeof(double), d * sizeof(double),
                            k, hipMemcpyHostToDevice, s));
  return SC_OK;
}

// One Lloyd step of the chosen form into w (labels, sums, counts, st->dist).
int lloyd_step(const sc_index* ix, KmWork& w, int k, double* cent, bool mixed, bool update,
               hipStream_t s) {
  if (mixed) {
    if (int rc = kmeans_mixed_step_launch(ix, w, k, cent, w.state, s)) return rc;
  } else {
    if (int rc = kmeans_assign_launch(ix, ix->rows, static_cast<int64_t>(ix->n), cent, k, w.labels,
                                      w.state, s))
      return rc;
    if (int rc = kmeans_partial_launch(ix, w, k, w.state, s)) return rc;
  }
  return kmeans_finish_launch(ix, w, k, w.state, cent, update, s);
}

int kmeans_run_impl(const sc_index* index, int k, size_t max_iter, bool mixed, double* centroids,
                    uint32_t* labels, uint64_t* size, double* distortion, uint32_t* iters,
                    hipStream_t s, uint64_t* fallback) {
  KmWork w;
  if (int rc = w.alloc(index, k, s)) return rc;
  DevBuf cent;
  SC_HIP(cent.alloc(sizeof(double) * k * index->ld, s));
  if (int rc = upload_centroids(centroids, k, index->d, index->ld, cent.as<double>(), s)) return rc;
  if (int rc = kmeans_state_init(w.state, s)) return rc;
  if (int rc = kmeans_x2_launch(index, w, s)) return rc;
  for (size_t it = 0; it < max_iter; ++it) {
    if (int rc = lloyd_step(index, w, k, cent.as<double>(), mixed, true, s)) return rc;
    if (int rc = kmeans_control_launch(w.state, s)) return rc;
  }
  KmState hs{};
  if (index->d > 0)
    SC_HIP(hipMemcpy2DAsync(centroids, index->d * sizeof(double), cent.p,
                            index->ld * sizeof(double), index->d * sizeof(double), k,
                            hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(labels, w.labels, sizeof(uint32_t) * index->n, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(size, w.counts, sizeof(uint64_t) * k, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(&hs, w.state, sizeof(KmState), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  *distortion = hs.distortion;
  *iters = hs.iters;
  if (fallback) *fallback = hs.fallback;
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_kmeans_step(const sc_index* index, const double* centroids, size_t k, double* sums,
                   uint64_t* counts, uint32_t* membership, double* distortion) {
  clear_error();
  if (!index || !centroids || !sums || !counts || !membership || !distortion)
    return fail(SC_PARAMETERS_ERROR, "null argument");
  if (k < 1) return fail(SC_PARAMETERS_ERROR, "k must be >= 1");
  if (index->n == 0) {
    for (size_t i = 0; i < k; ++i) counts[i] = 0;
    for (size_t i = 0; i < k * index->d; ++i) sums[i] = 0.0;
    *distortion = 0.0;
    return SC_OK;
  }
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  KmWork w;
  if (int rc = w.alloc(index, static_cast<int>(k), s)) return rc;
  DevBuf cent;
  SC_HIP(cent.alloc(sizeof(double) * k * index->ld, s));
  if (int rc = upload_centroids(centroids, static_cast<int>(k), index->d, index->ld,
                                cent.as<double>(), s))
    return rc;
  if (int rc = kmeans_state_init(w.state, s)) return rc;
  if (int rc = kmeans_x2_launch(index, w, s)) return rc;
  if (int rc = lloyd_step(index, w, static_cast<int>(k), cent.as<double>(), false, false, s))
    return rc;
  KmState hs{};
  if (index->d > 0)
    SC_HIP(hipMemcpyAsync(sums, w.sums, sizeof(double) * k * index->d, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(counts, w.counts, sizeof(uint64_t) * k, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(membership, w.labels, sizeof(uint32_t) * index->n, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(&hs, w.state, sizeof(KmState), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  *distortion = hs.dist;
  return SC_OK;
}

int sc_kmeans_run(const sc_index* index, size_t k, size_t max_iter, sc_mode mode,
                  double* centroids, uint32_t* labels, uint64_t* size, double* distortion,
                  uint32_t* iters) {
  clear_error();
  // kmeans.rs:257-269
  if (k < 2) return fail(SC_FIT_FAILED, "invalid number of clusters: " + std::to_string(k));
  if (max_iter == 0)
    return fail(SC_FIT_FAILED, "invalid maximum number of iterations: " + std::to_string(max_iter));
  if (!index || !centroids || !labels || !size || !distortion || !iters)
    return fail(SC_PARAMETERS_ERROR, "null argument");
  if (mode != SC_PLAIN_FP64 && mode != SC_MIXED) return fail(SC_ERR_UNSUPPORTED, "unknown mode");
  if (index->n == 0) return fail(SC_FIT_FAILED, "no rows");
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  const bool mixed = mode == SC_MIXED && kmeans_mixed_supported(index, static_cast<int>(k));
  uint64_t fb = 0;
  const int rc = kmeans_run_impl(index, static_cast<int>(k), max_iter, mixed, centroids, labels,
                                 size, distortion, iters, s, &fb);
  sc_stats& st = last_stats();
  st = sc_stats{};
  st.queries = index->n;
  st.uncertified = fb;
  return rc;
}

int sc_kmeans_predict(const void* x, sc_dtype dtype, size_t n, size_t d, const double* centroids,
                      size_t k, uint32_t* labels) {
  clear_error();
  if (!centroids || !labels || (!x && n > 0)) return fail(SC_PARAMETERS_ERROR, "null argument");
  if (dtype != SC_F64 && dtype != SC_F32) return fail(SC_ERR_UNSUPPORTED, "dtype must be f64 or f32");
  if (k < 1) return fail(SC_PARAMETERS_ERROR, "k must be >= 1");
  if (n == 0) return SC_OK;
  int dev = 0;
  SC_HIP(hipGetDevice(&dev));
  hipStream_t s = stream_for_device(dev);
  sc_index tmp{};
  tmp.dtype = dtype;
  tmp.n = n;
  tmp.d = d;
  tmp.ld = padded_dim(d);
  tmp.device = dev;
  DevBuf rows, cent, lab;
  SC_HIP(rows.alloc(elem_size(dtype) * n * tmp.ld, s));
  SC_HIP(cent.alloc(sizeof(double) * k * tmp.ld, s));
  SC_HIP(lab.alloc(sizeof(uint32_t) * n, s));
  if (int rc = upload_padded(x, dtype, n, d, tmp.ld, rows.p, s)) return rc;
  if (int rc = upload_centroids(centroids, static_cast<int>(k), d, tmp.ld, cent.as<double>(), s))
    return rc;
  if (int rc = kmeans_assign_launch(&tmp, rows.p, static_cast<int64_t>(n), cent.as<double>(),
                                    static_cast<int>(k), lab.as<uint32_t>(), nullptr, s))
    return rc;
  SC_HIP(hipMemcpyAsync(labels, lab.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

int sc_kmeans_plus_plus(const sc_index* index, size_t k, int has_seed, uint64_t seed,
                        uint32_t* labels) {
  clear_error();
  if (!index || !labels) return fail(SC_PARAMETERS_ERROR, "null argument");
  const int64_t n = static_cast<int64_t>(index->n);
  if (n == 0) return fail(SC_PARAMETERS_ERROR, "no rows");
  if (k < 1) return fail(SC_PARAMETERS_ERROR, "k must be >= 1");
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  // rand_custom.rs:8-33: Some(seed) -> seed_from_u64(seed); None -> 0
  SmallRng rng(has_seed ? seed : 0);
  const size_t es = elem_size(index->dtype);
  const size_t row_bytes = index->ld * es;
  DevBuf cen, dmin, dy;
  SC_HIP(cen.alloc(row_bytes, s));
  SC_HIP(dmin.alloc(sizeof(double) * n, s));
  SC_HIP(dy.alloc(sizeof(uint32_t) * n, s));
  SC_HIP(hipMemsetAsync(dy.p, 0, sizeof(uint32_t) * n, s));
  const unsigned grid = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, s, dmin.as<double>(), n, DBL_MAX);
  SC_HIP(hipGetLastError());
  const char* base = static_cast<const char*>(index->rows);
  // kmeans.rs:358-362: first centre is a uniformly drawn row
  const uint64_t pick = rng.gen_range(static_cast<uint64_t>(n));
  SC_HIP(hipMemcpyAsync(cen.p, base + pick * row_bytes, row_bytes, hipMemcpyDeviceToDevice, s));
  std::vector<double> hd(static_cast<size_t>(n));
  auto round = [&](uint32_t label) -> int {
    if (index->dtype == SC_F64)
      hipLaunchKernelGGL(kmeanspp_round_kernel<double>, dim3(grid), dim3(256), 0, s,
                         static_cast<const double*>(index->rows), n, static_cast<int>(index->ld),
                         cen.as<double>(), dmin.as<double>(), dy.as<uint32_t>(), label);
    else
      hipLaunchKernelGGL(kmeanspp_round_kernel<float>, dim3(grid), dim3(256), 0, s,
                         static_cast<const float*>(index->rows), n, static_cast<int>(index->ld),
                         cen.as<float>(), dmin.as<double>(), dy.as<uint32_t>(), label);
    SC_HIP(hipGetLastError());
    return SC_OK;
  };
  for (size_t j = 1; j < k; ++j) {
    if (int rc = round(static_cast<uint32_t>(j - 1))) return rc;
    SC_HIP(hipMemcpyAsync(hd.data(), dmin.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    // kmeans.rs:381-394: sequential sum, cutoff, linear cumulative scan
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum += hd[i];
    const double cutoff = rng.gen_f64() * sum;
    double cost = 0.0;
    int64_t index_i = 0;
    while (index_i < n) {
      cost += hd[index_i];
      if (cost >= cutoff) break;
      index_i += 1;
    }
    // the reference would index past the end here (only reachable with a
    // NaN sum); stay on the last row instead
    if (index_i >= n) index_i = n - 1;
    SC_HIP(hipMemcpyAsync(cen.p, base + index_i * row_bytes, row_bytes, hipMemcpyDeviceToDevice, s));
  }
  if (int rc = round(static_cast<uint32_t>(k - 1))) return rc;
  SC_HIP(hipMemcpyAsync(labels, dy.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

}  // extern "C"
