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
//   reduce    Two kernels shared by every form and by the batched form.
//             group: block sums added in groups of 16 blocks, in block
//             order.  final: one workgroup per cluster adds the groups in
//             order, computes the cluster's distortion term and (update)
//             centroid_i = sums_i / size_i when size_i > 0
//             (kmeans.rs:297-303); the last workgroup of a problem adds the
//             terms into the distortion
//               D = sum_i ||x_i - mu||^2 - 2 sum_c (c - mu).(S_c - n_c mu)
//                   + sum_c n_c ||c - mu||^2
//             (the cost BBDTree::clustering returns, bbd_tree.rs:162 and
//             :297-305, is the same quantity, sum_i ||x_i - c_i||^2; the
//             shift mu, the first row, keeps the three sums near the data's
//             own spread, so rows far from the origin do not turn D into a
//             difference of large numbers) and runs the stop rule
//             `if distortion <= dist break` (kmeans.rs:305-309) on the
//             device; later kernels return at once, so a fit needs one host
//             synchronisation.
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

// The shift mu of the distortion sums: the first row, widened to f64, with
// non-finite values replaced by 0.
template <typename T>
__global__ void mu_kernel(const T* __restrict__ rows, int ld, double* __restrict__ mu) {
  for (int j = threadIdx.x; j < ld; j += blockDim.x) {
    const double v = static_cast<double>(rows[j]);
    mu[j] = isfinite(v) ? v : 0.0;
  }
}

// Per block: sum of ||x - mu||^2 over its rows (each row a sequential f64
// fold, 8 consecutive rows per thread, then a fixed tree).
template <typename T>
__global__ __launch_bounds__(NT) void x2_block_kernel(const T* __restrict__ rows, int64_t n, int ld,
                                                      const double* __restrict__ mu,
                                                      double* __restrict__ x2block) {
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kBlockRows;
  double s = 0.0;
  for (int j = 0; j < kBlockRows / NT; ++j) {
    const int64_t r = r0 + tid * (kBlockRows / NT) + j;
    if (r < n) s = s + row_shift_x2(rows + r * ld, mu, ld);
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

// Group sums.  Item (problem, group * slices + slice): for 256 consecutive
// elements e of k x d, gsum[g][e] = psum over blocks 16g .. 16g+15 added in
// block order (and the counts of the clusters whose first element is in the
// slice).  All sixteen loads are issued before the adds.
__global__ __launch_bounds__(kReduceChunk) void kmeans_group_kernel(
    const KmProb* __restrict__ P, const uint2* __restrict__ items) {
  const uint2 it = items[blockIdx.x];
  const KmProb p = P[it.x];
  const int64_t kd = static_cast<int64_t>(p.k) * p.d;
  const int64_t slices = (kd + kReduceChunk - 1) / kReduceChunk;
  const int64_t g = it.y / slices;
  const int64_t e = (it.y % slices) * kReduceChunk + threadIdx.x;
  if (e >= kd) return;
  const int64_t b0 = g * kGroupBlocks;
  const int nbk = static_cast<int>(min<int64_t>(kGroupBlocks, p.nb - b0));
  // the partials are loaded before the stop flag is read, so both requests
  // are in flight together; every load is unconditional (a block index
  // clamped into range; values past the group's end are not added), so
  // that all of them are issued before the first add
  double v[kGroupBlocks];
#pragma unroll
  for (int i = 0; i < kGroupBlocks; ++i)
    v[i] = to_global(p.psum)[min<int64_t>(b0 + i, p.nb - 1) * kd + e];
  if (*reinterpret_cast<volatile const uint32_t*>(&p.st->done)) return;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < kGroupBlocks; ++i)
    if (i < nbk) s = s + v[i];
  to_global(p.gsum)[g * kd + e] = s;
  if (e % p.d == 0) {
    // the count: integer adds, loaded together (any order gives the same sum)
    const int64_t c = e / p.d;
    uint32_t w[kGroupBlocks];
#pragma unroll
    for (int i = 0; i < kGroupBlocks; ++i)
      w[i] = to_global(p.pcount)[min<int64_t>(b0 + i, p.nb - 1) * p.k + c];
    uint64_t cnt = 0;
#pragma unroll
    for (int i = 0; i < kGroupBlocks; ++i)
      if (i < nbk) cnt += w[i];
    to_global(p.gcount)[g * p.k + c] = cnt;
  }
}

__device__ __forceinline__ uint16_t bf16_rne_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Three-term BF16 split of v = -2c (c rounded to f32 first): hi = RN(v),
// mid = RN(v - hi), lo = RN(v - hi - mid).  Both differences are exact in
// f32, so |v - hi - mid - lo| <= 2^-24 |v|.
__device__ __forceinline__ void split_m2(double c, uint16_t& hi, uint16_t& mid, uint16_t& lo) {
  const float v = -2.0f * static_cast<float>(c);
  hi = bf16_rne_bits(v);
  const float e1 = v - __uint_as_float(static_cast<uint32_t>(hi) << 16);
  mid = bf16_rne_bits(e1);
  lo = bf16_rne_bits(e1 - __uint_as_float(static_cast<uint32_t>(mid) << 16));
}

// Largest finite ||c||^2 of a problem, rounded up, and its square root.
__device__ __forceinline__ void set_cmax_of(const KmProb& p, double mx) {
  p.cmax[0] = __double2float_ru(sqrt(mx) * (1.0 + 9.5367431640625e-07));
  p.cmax[1] = __double2float_ru(mx * (1.0 + 9.5367431640625e-07));
}

__device__ __forceinline__ void set_cmax(const KmProb& p) {
  const volatile double* c2 = p.c2d;
  double mx = 0.0;
  for (int c = 0; c < p.k; ++c) {
    const double v = c2[c];
    if (v < 1e300) mx = fmax(mx, v);  // NaN and inf centroids never count
  }
  set_cmax_of(p, mx);
}

// Final sums.  Item (problem, cluster c): sums_c = group sums in group
// order, the count, the distortion term of c with this step's centroid (dot
// and ||c||^2 added in dimension order), then (update) c = sums_c / n_c when
// n_c > 0, and the MFMA terms of the resulting centroid when the problem has
// them.  LAST: the last workgroup of a problem (an arrival counter) adds the
// terms in cluster order, applies the stop rule (control) and sets the
// centroid-norm maxima; otherwise the close kernel does that after this
// launch.  Each arrival takes a device-scope release fence, which writes
// back the L2: cheap for one problem's k workgroups, not for the tens of
// thousands of a batched call.
// FUSED (one problem, ng * d <= kFusedGroupSums): the workgroup first forms
// the group sums of its cluster itself, from the block partials (each a sum
// of kGroupBlocks blocks in block order, into LDS), which is the arithmetic
// of kmeans_group_kernel, so the group launch is left out.
constexpr int kFusedThreads = 512;
constexpr int kFusedGroupSums = 8192;  // doubles of LDS for the group sums

template <bool LAST, bool FUSED>
__global__ __launch_bounds__(FUSED ? kFusedThreads : 128) void kmeans_final_kernel(
    const KmProb* __restrict__ P, const uint2* __restrict__ items, int update, int control) {
  constexpr int NTH = FUSED ? kFusedThreads : 128;
  const uint2 it = items[blockIdx.x];
  const KmProb p = P[it.x];
  if (p.st->done) return;
  const int c = static_cast<int>(it.y);
  const int tid = threadIdx.x;
  const int d = p.d;
  const int64_t kd = static_cast<int64_t>(p.k) * d;
  const bool tables = p.bhi != nullptr;
  __shared__ double s_dot[128], s_c2[128], s_n2[128];
  __shared__ double s_gs[FUSED ? kFusedGroupSums : 1];  // [group][dimension]
  __shared__ uint64_t s_n;
  __shared__ int s_last;
  // the count: integer adds, so any order gives the same value
  if (tid == 0) s_n = 0;
  __syncthreads();
  {
    uint64_t n = 0;
    if (FUSED) {
      for (int64_t b = tid; b < p.nb; b += NTH) n += to_global(p.pcount)[b * p.k + c];
    } else {
      for (int64_t g = tid; g < p.ng; g += NTH) n += to_global(p.gcount)[g * p.k + c];
    }
    if (n) atomicAdd(reinterpret_cast<unsigned long long*>(&s_n), static_cast<unsigned long long>(n));
  }
  if (FUSED) {
    // group sums of this cluster: (group, dimension) pairs, dimensions
    // across consecutive threads; each adds its group's blocks in order.
    // A thread requests the partials of four pairs before adding any.
    constexpr int U = 4;
    const auto ps = to_global(p.psum) + static_cast<int64_t>(c) * d;
    const int64_t np = p.ng * d;
    for (int64_t e0 = tid; e0 < np; e0 += static_cast<int64_t>(NTH) * U) {
      double v[U][kGroupBlocks];
      int nbk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t e = e0 + static_cast<int64_t>(u) * NTH;
        const int64_t g = e / d;
        const int j = static_cast<int>(e - g * d);
        const int64_t b0 = g * kGroupBlocks;
        nbk[u] = e < np ? static_cast<int>(min<int64_t>(kGroupBlocks, p.nb - b0)) : 0;
        // unconditional loads (block index clamped into range; values past
        // the group's end are not added), all issued before the first add
#pragma unroll
        for (int i = 0; i < kGroupBlocks; ++i) v[u][i] = ps[min<int64_t>(b0 + i, p.nb - 1) * kd + j];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        double sg = 0.0;
#pragma unroll
        for (int i = 0; i < kGroupBlocks; ++i)
          if (i < nbk[u]) sg = sg + v[u][i];
        if (nbk[u] > 0) s_gs[e0 + static_cast<int64_t>(u) * NTH] = sg;
      }
    }
  }
  __syncthreads();
  const uint64_t n = s_n;
  if (tid == 0) to_global(p.counts)[c] = n;
  const bool upd = update && n > 0;
  double dot = 0.0, c2 = 0.0, n2 = 0.0;
  for (int j0 = 0; j0 < d; j0 += 128) {
    const int j = j0 + tid;
    if (tid < 128 && j < d) {
      double S = 0.0;
      if (FUSED) {
        // group sums in group order
        for (int64_t g = 0; g < p.ng; ++g) S = S + s_gs[g * d + j];
      } else {
        const auto gs = to_global(p.gsum) + static_cast<int64_t>(c) * d + j;
        // group sums in group order, 32 requested before they are added
        // (unconditional loads, the group index clamped into range)
        for (int64_t g0 = 0; g0 < p.ng; g0 += 32) {
          double v[32];
#pragma unroll
          for (int i = 0; i < 32; ++i) v[i] = gs[min<int64_t>(g0 + i, p.ng - 1) * kd];
#pragma unroll
          for (int i = 0; i < 32; ++i)
            if (g0 + i < p.ng) S = S + v[i];
        }
      }
      to_global(p.sums)[static_cast<int64_t>(c) * d + j] = S;
      const double cj = to_global(p.cent)[static_cast<int64_t>(c) * p.ld + j];
      const double nc = upd ? S / static_cast<double>(n) : cj;
      const double mj = to_global(p.mu)[j];
      const double dc = cj - mj;
      const double dS = S - static_cast<double>(n) * mj;
      s_dot[tid] = dc * dS;
      s_c2[tid] = dc * dc;
      s_n2[tid] = nc * nc;
      if (upd) to_global(p.cent)[static_cast<int64_t>(c) * p.ld + j] = nc;
      if (tables) {
        uint16_t h, m, l;
        split_m2(nc, h, m, l);
        to_global(p.bhi)[static_cast<int64_t>(c) * p.ld + j] = h;
        to_global(p.bmd)[static_cast<int64_t>(c) * p.ld + j] = m;
        to_global(p.blo)[static_cast<int64_t>(c) * p.ld + j] = l;
      }
    }
    __syncthreads();
    // the three sums in dimension order, one lane each
    if (tid < 3) {
      const double* arr = tid == 0 ? s_dot : (tid == 1 ? s_c2 : s_n2);
      double a = tid == 0 ? dot : (tid == 1 ? c2 : n2);
      const int m = min(128, d - j0);
      for (int jj = 0; jj < m; ++jj) a = a + arr[jj];
      if (tid == 0) dot = a;
      if (tid == 1) c2 = a;
      if (tid == 2) n2 = a;
    }
    __syncthreads();
  }
  if (tid == 1) s_c2[0] = c2;
  if (tid == 2) s_n2[0] = n2;
  __syncthreads();
  c2 = s_c2[0];
  n2 = s_n2[0];
  __syncthreads();
  if (tid == 0) {
    to_global(p.term)[c] = n > 0 ? static_cast<double>(n) * c2 - 2.0 * dot : 0.0;
    if (tables) {
      to_global(p.c2f)[c] = static_cast<float>(n2);
      to_global(p.c2d)[c] = n2;
    }
  }
  if constexpr (!LAST) return;
  if (tid == 0) {
    __threadfence();
    const uint32_t old = atomicAdd(p.arrive, 1u);
    s_last = old == static_cast<uint32_t>(p.k - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  // The last workgroup: the terms (and centroid norms) of all clusters,
  // loaded by 128 threads at a time and added in cluster order by one.
  double s = 0.0, mx = 0.0;
  if (tid == 0) s = p.st->x2total;
  for (int c0 = 0; c0 < p.k; c0 += 128) {
    const int cc = c0 + tid;
    if (tid < 128 && cc < p.k) {
      s_dot[tid] = __hip_atomic_load(p.term + cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (tables)
        s_c2[tid] = __hip_atomic_load(p.c2d + cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (tid == 0) {
      const int m = min(128, p.k - c0);
      for (int j = 0; j < m; ++j) s = s + s_dot[j];
      if (tables)
        for (int j = 0; j < m; ++j)
          if (s_c2[j] < 1e300) mx = fmax(mx, s_c2[j]);  // NaN and inf centroids never count
    }
    __syncthreads();
  }
  if (tid != 0) return;
  p.st->dist = s;
  if (control) {
    p.st->iters += 1;
    if (p.st->distortion <= s) {
      p.st->done = 1;
    } else {
      p.st->distortion = s;
    }
  }
  if (tables) set_cmax_of(p, mx);
  *p.arrive = 0u;
}

// One workgroup per problem, after kmeans_final_kernel<false>: the
// distortion (the terms of all clusters, loaded by every thread 128 at a
// time and added in cluster order by one, as the last workgroup of the
// arrival form does), the stop rule (control) and the centroid-norm maxima.
__global__ __launch_bounds__(128) void kmeans_close_kernel(const KmProb* __restrict__ P, int control) {
  const KmProb p = P[blockIdx.x];
  if (p.st->done) return;
  const int tid = threadIdx.x;
  const bool tables = p.bhi != nullptr;
  __shared__ double s_term[128], s_c2[128];
  double s = 0.0, mx = 0.0;
  if (tid == 0) s = p.st->x2total;
  for (int c0 = 0; c0 < p.k; c0 += 128) {
    const int cc = c0 + tid;
    if (cc < p.k) {
      s_term[tid] = to_global(p.term)[cc];
      if (tables) s_c2[tid] = to_global(p.c2d)[cc];
    }
    __syncthreads();
    if (tid == 0) {
      const int m = min(128, p.k - c0);
      for (int j = 0; j < m; ++j) s = s + s_term[j];
      if (tables)
        for (int j = 0; j < m; ++j)
          if (s_c2[j] < 1e300) mx = fmax(mx, s_c2[j]);  // NaN and inf centroids never count
    }
    __syncthreads();
  }
  if (tid != 0) return;
  p.st->dist = s;
  if (control) {
    p.st->iters += 1;
    if (p.st->distortion <= s) {
      p.st->done = 1;
    } else {
      p.st->distortion = s;
    }
  }
  if (tables) set_cmax_of(p, mx);
}

// MFMA terms of the current centroids: one workgroup per problem.  Padding
// rows (k .. kpad) and padding dimensions get zeros and c2f = +inf.
__global__ __launch_bounds__(NT) void kmeans_tables_kernel(const KmProb* __restrict__ P) {
  const KmProb p = P[blockIdx.x];
  if (p.bhi == nullptr) return;
  for (int64_t e = threadIdx.x; e < static_cast<int64_t>(p.kpad) * p.ld; e += NT) {
    const int c = static_cast<int>(e / p.ld), j = static_cast<int>(e % p.ld);
    uint16_t h = 0, m = 0, l = 0;
    if (c < p.k && j < p.d) split_m2(p.cent[static_cast<int64_t>(c) * p.ld + j], h, m, l);
    p.bhi[e] = h;
    p.bmd[e] = m;
    p.blo[e] = l;
  }
  for (int c = threadIdx.x; c < p.kpad; c += NT) {
    if (c < p.k) {
      const double* cc = p.cent + static_cast<int64_t>(c) * p.ld;
      double s = 0.0;
      for (int j = 0; j < p.d; ++j) s = s + cc[j] * cc[j];
      p.c2f[c] = static_cast<float>(s);
      p.c2d[c] = s;
    } else {
      p.c2f[c] = INFINITY;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    set_cmax(p);
    *p.arrive = 0u;
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
  const int ld = static_cast<int>(ix->ld);
  if (n > 0) {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(mu_kernel<double>, dim3(1), dim3(NT), 0, s,
                         static_cast<const double*>(ix->rows), ld, w.p.mu);
    else
      hipLaunchKernelGGL(mu_kernel<float>, dim3(1), dim3(NT), 0, s,
                         static_cast<const float*>(ix->rows), ld, w.p.mu);
  } else {
    hipLaunchKernelGGL(fill_kernel, dim3(static_cast<unsigned>((ld + NT - 1) / NT)), dim3(NT), 0,
                       s, w.p.mu, static_cast<int64_t>(ld), 0.0);
  }
  SC_HIP(hipGetLastError());
  if (w.p.nb > 0) {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(x2_block_kernel<double>, dim3(static_cast<unsigned>(w.p.nb)), dim3(NT),
                         0, s, static_cast<const double*>(ix->rows), n, ld, w.p.mu, w.p.x2block);
    else
      hipLaunchKernelGGL(x2_block_kernel<float>, dim3(static_cast<unsigned>(w.p.nb)), dim3(NT),
                         0, s, static_cast<const float*>(ix->rows), n, ld, w.p.mu, w.p.x2block);
    SC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(x2_total_kernel, dim3(1), dim3(1), 0, s, w.p.x2block, w.p.nb, w.p.st);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int kmeans_partial_launch(const sc_index* ix, KmWork& w, int k, const KmState* st, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  const int d = static_cast<int>(ix->d);
  const int ld = static_cast<int>(ix->ld);
  if (w.p.nb == 0) return SC_OK;
  ProfScope ps("kmeans_partial", s);
  const size_t lds = partial_lds_bytes(ix, k);
  const dim3 grid(static_cast<unsigned>(w.p.nb));
  if (static_cast<int64_t>(k) * d <= kMaxLdsAcc && lds <= 96 * 1024) {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(kmeans_partial_lds_kernel<double>, grid, dim3(NT), lds, s,
                         static_cast<const double*>(ix->rows), n, ld, d, w.p.labels, k, w.p.psum,
                         w.p.pcount, st);
    else
      hipLaunchKernelGGL(kmeans_partial_lds_kernel<float>, grid, dim3(NT), lds, s,
                         static_cast<const float*>(ix->rows), n, ld, d, w.p.labels, k, w.p.psum,
                         w.p.pcount, st);
  } else {
    if (ix->dtype == SC_F64)
      hipLaunchKernelGGL(kmeans_partial_ballot_kernel<double>, grid, dim3(NT), 0, s,
                         static_cast<const double*>(ix->rows), n, ld, d, w.p.labels, k, w.p.psum,
                         w.p.pcount, st);
    else
      hipLaunchKernelGGL(kmeans_partial_ballot_kernel<float>, grid, dim3(NT), 0, s,
                         static_cast<const float*>(ix->rows), n, ld, d, w.p.labels, k, w.p.psum,
                         w.p.pcount, st);
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int kmeans_reduce_launch(const KmProb* d_probs, const KmItems& items, bool update, bool control,
                         hipStream_t s) {
  ProfScope ps("kmeans_reduce", s);
  if (items.fused && items.nprob == 1) {
    // one problem whose group sums fit LDS: the final kernel forms them
    if (items.nfinal > 0) {
      hipLaunchKernelGGL((kmeans_final_kernel<true, true>), dim3(static_cast<unsigned>(items.nfinal)),
                         dim3(kFusedThreads), 0, s, d_probs, items.fin, update ? 1 : 0,
                         control ? 1 : 0);
      SC_HIP(hipGetLastError());
    }
    return SC_OK;
  }
  if (items.ngroup > 0) {
    hipLaunchKernelGGL(kmeans_group_kernel, dim3(static_cast<unsigned>(items.ngroup)),
                       dim3(kReduceChunk), 0, s, d_probs, items.group);
    SC_HIP(hipGetLastError());
  }
  if (items.nfinal > 0) {
    if (items.nprob <= 1) {
      hipLaunchKernelGGL((kmeans_final_kernel<true, false>), dim3(static_cast<unsigned>(items.nfinal)),
                         dim3(128), 0, s, d_probs, items.fin, update ? 1 : 0, control ? 1 : 0);
    } else {
      hipLaunchKernelGGL((kmeans_final_kernel<false, false>), dim3(static_cast<unsigned>(items.nfinal)),
                         dim3(128), 0, s, d_probs, items.fin, update ? 1 : 0, control ? 1 : 0);
      SC_HIP(hipGetLastError());
      hipLaunchKernelGGL(kmeans_close_kernel, dim3(static_cast<unsigned>(items.nprob)), dim3(128),
                         0, s, d_probs, control ? 1 : 0);
    }
    SC_HIP(hipGetLastError());
  }
  return SC_OK;
}

int kmeans_tables_launch(const KmProb* d_probs, size_t count, hipStream_t s) {
  if (count == 0) return SC_OK;
  hipLaunchKernelGGL(kmeans_tables_kernel, dim3(static_cast<unsigned>(count)), dim3(NT), 0, s,
                     d_probs);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int KmItems::build(const KmProb* probs, size_t count, hipStream_t s) {
  std::vector<uint2> g, f;
  for (size_t i = 0; i < count; ++i) {
    const KmProb& p = probs[i];
    const int64_t kd = static_cast<int64_t>(p.k) * p.d;
    const int64_t slices = (kd + kReduceChunk - 1) / kReduceChunk;
    for (int64_t t = 0; t < p.ng * slices; ++t)
      g.push_back(make_uint2(static_cast<unsigned>(i), static_cast<unsigned>(t)));
    for (int c = 0; c < p.k; ++c) f.push_back(make_uint2(static_cast<unsigned>(i), static_cast<unsigned>(c)));
  }
  ngroup = static_cast<int64_t>(g.size());
  nfinal = static_cast<int64_t>(f.size());
  nprob = static_cast<int64_t>(count);
  fused = count == 1 && probs[0].ng * std::max(probs[0].d, 1) <= kFusedGroupSums;
  // staged copies: the host vectors may die when they return, and the host
  // does not wait for work already queued on s
  if (ngroup) {
    SC_HIP(b_group.alloc(sizeof(uint2) * ngroup, s));
    if (int rc = upload_staged(b_group.p, g.data(), sizeof(uint2) * ngroup, s)) return rc;
    group = b_group.as<uint2>();
  }
  if (nfinal) {
    SC_HIP(b_final.alloc(sizeof(uint2) * nfinal, s));
    if (int rc = upload_staged(b_final.p, f.data(), sizeof(uint2) * nfinal, s)) return rc;
    fin = b_final.as<uint2>();
  }
  return SC_OK;
}

int KmWork::alloc(const sc_index* ix, int k, const double* d_cent, bool mixed, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  const int64_t d = static_cast<int64_t>(ix->d);
  const int64_t d1 = std::max<int64_t>(d, 1);
  p = KmProb{};
  p.n = n;
  p.d = static_cast<int>(d);
  p.ld = static_cast<int>(ix->ld);
  p.k = k;
  p.kpad = (k + 31) / 32 * 32;
  p.nb = (n + kBlockRows - 1) / kBlockRows;
  p.ng = (p.nb + kGroupBlocks - 1) / kGroupBlocks;
  const int64_t nb1 = std::max<int64_t>(p.nb, 1), ng1 = std::max<int64_t>(p.ng, 1);
  SC_HIP(b_labels.alloc(sizeof(uint32_t) * std::max<int64_t>(n, 1), s));
  SC_HIP(b_psum.alloc(sizeof(double) * nb1 * k * d1, s));
  SC_HIP(b_pcount.alloc(sizeof(uint32_t) * nb1 * k, s));
  SC_HIP(b_gsum.alloc(sizeof(double) * ng1 * k * d1, s));
  SC_HIP(b_gcount.alloc(sizeof(uint64_t) * ng1 * k, s));
  SC_HIP(b_sums.alloc(sizeof(double) * k * d1, s));
  SC_HIP(b_counts.alloc(sizeof(uint64_t) * k, s));
  // x2 per block | k distortion terms | k centroid norms (MFMA form) | mu
  SC_HIP(b_x2.alloc(sizeof(double) * (nb1 + 2 * static_cast<int64_t>(k) + p.ld), s));
  // the state, then the arrival counter
  SC_HIP(b_state.alloc(sizeof(KmState) + 16, s));
  SC_HIP(b_prob.alloc(sizeof(KmProb), s));
  p.rows = ix->rows;
  p.cent = const_cast<double*>(d_cent);
  p.labels = b_labels.as<uint32_t>();
  p.psum = b_psum.as<double>();
  p.pcount = b_pcount.as<uint32_t>();
  p.gsum = b_gsum.as<double>();
  p.gcount = b_gcount.as<uint64_t>();
  p.sums = b_sums.as<double>();
  p.counts = b_counts.as<uint64_t>();
  p.x2block = b_x2.as<double>();
  p.term = p.x2block + nb1;
  p.mu = p.term + 2 * static_cast<int64_t>(k);
  p.st = b_state.as<KmState>();
  p.arrive = reinterpret_cast<uint32_t*>(b_state.as<char>() + sizeof(KmState));
  SC_HIP(hipMemsetAsync(p.arrive, 0, 4, s));
  if (mixed) {
    const int64_t tab = static_cast<int64_t>(p.kpad) * p.ld;
    SC_HIP(b_tab.alloc(sizeof(uint16_t) * 3 * tab + sizeof(float) * (p.kpad + 4), s));
    p.bhi = b_tab.as<uint16_t>();
    p.bmd = p.bhi + tab;
    p.blo = p.bmd + tab;
    p.c2f = reinterpret_cast<float*>(p.blo + tab);
    p.cmax = p.c2f + p.kpad;
    p.c2d = p.term + k;
  }
  SC_HIP(hipMemcpyAsync(b_prob.p, &p, sizeof(KmProb), hipMemcpyHostToDevice, s));
  dprob = b_prob.as<KmProb>();
  // builds the item lists and waits for the stream (p's copy included)
  return items.build(&p, 1, s);
}

int kmeans_state_init(KmState* st, hipStream_t s) {
  hipLaunchKernelGGL(kmeans_state_init_kernel, dim3(1), dim3(1), 0, s, st);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc

using namespace sc;

namespace {

int upload_centroids(const double* host, int k, size_t d, size_t ld, double* dev, hipStream_t s) {
  if (d == 0) {
    SC_HIP(hipMemsetAsync(dev, 0, sizeof(double) * k * ld, s));
    return SC_OK;
  }
  return upload_staged_rows(dev, host, static_cast<size_t>(k), d * sizeof(double),
                            ld * sizeof(double), s);
}

// One Lloyd step of the chosen form into w (labels, sums, counts, st->dist;
// with `update` the new centroids, with `control` the stop rule).
int lloyd_step(const sc_index* ix, KmWork& w, bool mixed, bool update, bool control,
               hipStream_t s) {
  if (mixed) {
    if (int rc = kmeans_mixed_step_launch(ix, w, s)) return rc;
  } else {
    if (int rc = kmeans_assign_launch(ix, ix->rows, static_cast<int64_t>(ix->n), w.p.cent, w.p.k,
                                      w.p.labels, w.p.st, s))
      return rc;
    if (int rc = kmeans_partial_launch(ix, w, w.p.k, w.p.st, s)) return rc;
  }
  return kmeans_reduce_launch(w.dprob, w.items, update, control, s);
}

int kmeans_run_impl(const sc_index* index, int k, size_t max_iter, bool mixed, double* centroids,
                    uint32_t* labels, uint64_t* size, double* distortion, uint32_t* iters,
                    hipStream_t s, uint64_t* fallback) {
  DevBuf cent;
  SC_HIP(cent.alloc(sizeof(double) * k * index->ld, s));
  if (int rc = upload_centroids(centroids, k, index->d, index->ld, cent.as<double>(), s)) return rc;
  KmWork w;
  if (int rc = w.alloc(index, k, cent.as<double>(), mixed, s)) return rc;
  if (int rc = kmeans_state_init(w.p.st, s)) return rc;
  if (int rc = kmeans_x2_launch(index, w, s)) return rc;
  if (mixed)
    if (int rc = kmeans_tables_launch(w.dprob, 1, s)) return rc;
  // The max_iter steps (each a step launch and two reduce launches, all
  // returning at once after the stop rule fires) are captured in one hipGraph
  // and launched once: the graph removes the per-launch gaps between the
  // dependent kernels.  With the tracing hooks on, the steps are launched
  // one by one instead, so that every launch is timed.
  if (profiling_enabled()) {
    for (size_t it = 0; it < max_iter; ++it)
      if (int rc = lloyd_step(index, w, mixed, true, true, s)) return rc;
  } else {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int crc = SC_OK;
    for (size_t it = 0; it < max_iter && crc == SC_OK; ++it)
      crc = lloyd_step(index, w, mixed, true, true, s);
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (crc) {
      if (graph) (void)hipGraphDestroy(graph);
      return crc;
    }
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ce != hipSuccess) return hip_fail(ce, "hipGraphInstantiate");
    ce = hipGraphLaunch(exec, s);
    const hipError_t se = hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
    if (ce != hipSuccess) return hip_fail(ce, "hipGraphLaunch");
    if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize");
  }
  KmState hs{};
  if (index->d > 0)
    SC_HIP(hipMemcpy2DAsync(centroids, index->d * sizeof(double), cent.p,
                            index->ld * sizeof(double), index->d * sizeof(double), k,
                            hipMemcpyDeviceToHost, s));
  if (int rc = download_staged(labels, w.p.labels, sizeof(uint32_t) * index->n, s)) return rc;
  SC_HIP(hipMemcpyAsync(size, w.p.counts, sizeof(uint64_t) * k, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(&hs, w.p.st, sizeof(KmState), hipMemcpyDeviceToHost, s));
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
  DevBuf cent;
  SC_HIP(cent.alloc(sizeof(double) * k * index->ld, s));
  if (int rc = upload_centroids(centroids, static_cast<int>(k), index->d, index->ld,
                                cent.as<double>(), s))
    return rc;
  KmWork w;
  if (int rc = w.alloc(index, static_cast<int>(k), cent.as<double>(), false, s)) return rc;
  if (int rc = kmeans_state_init(w.p.st, s)) return rc;
  if (int rc = kmeans_x2_launch(index, w, s)) return rc;
  if (int rc = lloyd_step(index, w, false, false, false, s)) return rc;
  KmState hs{};
  if (index->d > 0)
    SC_HIP(hipMemcpyAsync(sums, w.p.sums, sizeof(double) * k * index->d, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(counts, w.p.counts, sizeof(uint64_t) * k, hipMemcpyDeviceToHost, s));
  if (int rc = download_staged(membership, w.p.labels, sizeof(uint32_t) * index->n, s)) return rc;
  SC_HIP(hipMemcpyAsync(&hs, w.p.st, sizeof(KmState), hipMemcpyDeviceToHost, s));
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
  if (int rc = download_staged(labels, lab.p, sizeof(uint32_t) * n, s)) return rc;
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
