// Exact brute-force kNN (LinearKNNSearch::find, linear_search.rs:52-84) for
// the Euclidian, Manhattan and Minkowski metrics on f64 / f32 rows.
//
// Three kernels:
//   scan    Tiled distance kernel.  A workgroup owns 32 queries x one
//           contiguous segment of the reference rows and sweeps the segment
//           in index order, 128 rows per tile, staging dimension chunks of
//           queries and rows through LDS.  Each thread folds a 4x4 register
//           micro-tile of sums in the reference's element order.  Per query
//           the workgroup keeps the k smallest sums seen so far in the
//           segment; a row whose sum is below that running k-th sum is
//           appended (index, sum) to the query's candidate list for the
//           segment.  Those lists are a superset of the rows that enter the
//           reference heap (an entrant's distance is below the global running
//           k-th distance, which never exceeds the segment's).
//   replay  One thread per query replays HeapSelection over the candidate
//           lists in segment order, i.e. in index order, exactly as
//           linear_search.rs:69-77 does over all rows: rows that would not
//           enter the heap are no-ops there too, so the heap array (order
//           included) is the reference's.
//   serial  For queries whose candidate list overflowed its capacity (or k
//           above the tiled limit): one workgroup per query computes 256
//           distances at a time and lane 0 replays the heap over those that
//           beat the current root.  Exact for any input, slower.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "metric.hpp"
#include "tile.hpp"

namespace sc {
namespace {

constexpr int QT = kTileQ;    // queries per workgroup
constexpr int RT = kTileR;    // reference rows per tile
constexpr int KMAX_TILED = 64;
constexpr int NT = kTileThreads;

static_assert(kTileD == kRowPad, "chunk must divide the row padding");

// HeapSelection::sift_down (heap_select.rs:71-86) on (dist, idx) arrays:
// children of kk are 2kk and 2kk+1; break on Equal or Greater.
__device__ inline void sift_down(double* hd, int64_t* hi, int kk, int n) {
  while (2 * kk <= n) {
    int j = 2 * kk;
    if (j < n && hd[j] < hd[j + 1]) j += 1;
    if (hd[kk] == hd[j] || hd[kk] > hd[j]) break;
    double td = hd[kk];
    hd[kk] = hd[j];
    hd[j] = td;
    int64_t ti = hi[kk];
    hi[kk] = hi[j];
    hi[j] = ti;
    kk = j;
  }
}

// Replacing the root of a valid heap and running heapify()
// (heap_select.rs:42-50) only moves elements in its final sift_down(0, n-1):
// every sift_down(i >= 1) meets a node already >= its children and breaks.
__device__ inline void heap_push_root(double* hd, int64_t* hi, int k, double d, int64_t idx) {
  hd[0] = d;
  hi[0] = idx;
  if (k > 1) sift_down(hd, hi, 0, k - 1);
}

template <typename T, int M>
__global__ __launch_bounds__(NT) void knn_scan_kernel(
    const T* __restrict__ refs, int64_t n, int ld, int d, const T* __restrict__ queries, int64_t q, int k,
    double p, int64_t seg_len, int nseg, uint32_t* __restrict__ cand_idx,
    double* __restrict__ cand_s, uint32_t* __restrict__ cand_count, int cap) {
  __shared__ TileSmem<T> sm;
  __shared__ double sbuf[QT][RT];
  __shared__ uint32_t mask[QT][RT / 32];
  __shared__ double topk[QT][KMAX_TILED];
  __shared__ double thr[QT];

  const int tid = threadIdx.x;
  const int tq = tid >> 5;  // 0..7 -> queries tq*4 .. tq*4+3
  const int tr = tid & 31;  // 0..31 -> rows tr + 32*j
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * QT;
  const int seg = blockIdx.y;
  const int64_t r_begin = static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);

  for (int e = tid; e < QT * KMAX_TILED; e += NT) (&topk[0][0])[e] = INFINITY;
  for (int e = tid; e < QT * (RT / 32); e += NT) (&mask[0][0])[e] = 0u;
  if (tid < QT) thr[tid] = INFINITY;

  uint32_t emitted = 0;  // candidates of query q0+tid (tid < QT)
  const int64_t my_q = q0 + tid;
  const size_t cand_base = (tid < QT && my_q < q)
                               ? (static_cast<size_t>(my_q) * nseg + seg) * static_cast<size_t>(cap)
                               : 0;

  for (int64_t t0 = r_begin; t0 < r_end; t0 += RT) {
    double acc[4][4];
    tile_sums<T, M>(sm, acc, refs, t0, r_end, queries, q0, q, ld, d, p);

    // Epilogue: mark rows whose sum beats the query's running k-th sum.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qi = tq * 4 + i;
      const double th = thr[qi];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int rl = tr + 32 * j;
        if (q0 + qi < q && t0 + rl < r_end && acc[i][j] < th) {
          sbuf[qi][rl] = acc[i][j];
          atomicOr(&mask[qi][rl >> 5], 1u << (rl & 31));
        }
      }
    }
    __syncthreads();
    if (tid < QT && my_q < q) {
      double* tk = topk[tid];
      for (int w = 0; w < RT / 32; ++w) {
        uint32_t bits = mask[tid][w];
        mask[tid][w] = 0u;
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1;
          const int rl = w * 32 + b;
          const double s = sbuf[tid][rl];
          if (s < tk[k - 1]) {
            int pos = k - 1;
            while (pos > 0 && tk[pos - 1] > s) {
              tk[pos] = tk[pos - 1];
              --pos;
            }
            tk[pos] = s;
            if (emitted < static_cast<uint32_t>(cap)) {
              cand_idx[cand_base + emitted] = static_cast<uint32_t>(t0 + rl);
              cand_s[cand_base + emitted] = s;
            }
            ++emitted;
          }
        }
      }
      thr[tid] = tk[k - 1];
    }
  }
  if (tid < QT && my_q < q) cand_count[static_cast<size_t>(my_q) * nseg + seg] = emitted;
}

// Manhattan on f32 rows with f32 sums (SC_MIXED).  Each term
// |a - b| is the reference's own f32 difference (manhattan.rs:56-59, the
// widening to f64 is exact), so only the summation differs: a sequential
// f32 sum A of the ld terms differs from their exact sum, and so from the
// reference's f64 fold S, by at most about (ld + 1) 2^-24 A.  With
// e = (ld + 4) 2^-24 (the extra terms cover the roundings of the bound
// arithmetic itself) and an absolute 2^-126 for underflow,
// L = A (1 - e) - 2^-126 <= S <= U = A (1 + e) + 2^-126.  A segment keeps
// the k smallest U it has seen (starting from `seed`, when given, the k-th
// smallest U of earlier rows, else +inf) and stores, in row order, every
// row whose L is below the current k-th of them: the k-th U is at or above
// the k-th S, so every row the reference heap admits is stored.  Their S is
// recomputed in the reference arithmetic (knn_rescore_kernel) and the
// usual replay runs over them.  A non-finite A (inf / NaN rows or an f32
// overflow) is always stored and never lowers the threshold.  With
// `out_thr` the segment's final k-th U is written per query (the seed of
// later segments).
__global__ __launch_bounds__(NT) void knn_scan_l1f32_kernel(
    const float* __restrict__ refs, int64_t r0, int64_t n, int ld,
    const float* __restrict__ queries, int64_t q, int k, int64_t seg_len, int nseg,
    const float* __restrict__ seed, float* __restrict__ out_thr,
    uint32_t* __restrict__ cand_idx, uint32_t* __restrict__ cand_count, int cap) {
  __shared__ float Qs[kTileD][QT + 1];
  __shared__ float Rs[kTileD][RT + 1];
  __shared__ uint32_t mask[QT][RT / 32];
  __shared__ float sL[QT][RT];
  __shared__ float sU[QT][RT];
  __shared__ float topk[QT][KMAX_TILED];
  __shared__ float thr[QT];

  const int tid = threadIdx.x;
  const int tq = tid >> 5;
  const int tr = tid & 31;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * QT;
  const int seg = blockIdx.y;
  const int64_t r_begin = r0 + static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);
  const float e = static_cast<float>(ld + 4) * 5.9604644775390625e-08f;
  const float tiny = 1.1754943508222875e-38f;

  for (int i = tid; i < QT * KMAX_TILED; i += NT) (&topk[0][0])[i] = INFINITY;
  for (int i = tid; i < QT * (RT / 32); i += NT) (&mask[0][0])[i] = 0u;
  if (tid < QT) thr[tid] = (seed && q0 + tid < q) ? seed[q0 + tid] : INFINITY;
  __syncthreads();
  // a seed is an upper bound on the k-th S: it caps every slot
  if (seed && tid < QT)
    for (int j = 0; j < k; ++j) topk[tid][j] = thr[tid];

  uint32_t emitted = 0;
  const int64_t my_q = q0 + tid;
  const size_t cand_base = (tid < QT && my_q < q)
                               ? (static_cast<size_t>(my_q) * nseg + seg) * static_cast<size_t>(cap)
                               : 0;

  for (int64_t t0 = r_begin; t0 < r_end; t0 += RT) {
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int c0 = 0; c0 < ld; c0 += kTileD) {
      __syncthreads();
      for (int x = tid; x < RT * kTileD; x += NT) {
        const int r = x / kTileD, c = x % kTileD;
        const int64_t row = t0 + r;
        Rs[c][r] = row < r_end ? refs[row * ld + c0 + c] : 0.0f;
      }
      for (int x = tid; x < QT * kTileD; x += NT) {
        const int r = x / kTileD, c = x % kTileD;
        const int64_t row = q0 + r;
        Qs[c][r] = row < q ? queries[row * ld + c0 + c] : 0.0f;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < kTileD; ++kk) {
        float a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Qs[kk][tq * 4 + i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Rs[kk][tr + 32 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = acc[i][j] + fabsf(a[i] - b[j]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qi = tq * 4 + i;
      const float th = thr[qi];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int rl = tr + 32 * j;
        const float A = acc[i][j];
        const bool fin = A < INFINITY;
        const float L = fin ? (A - A * e) - tiny : -INFINITY;
        if (q0 + qi < q && t0 + rl < r_end && L < th) {
          sL[qi][rl] = L;
          sU[qi][rl] = fin ? (A + A * e) + tiny : INFINITY;
          atomicOr(&mask[qi][rl >> 5], 1u << (rl & 31));
        }
      }
    }
    __syncthreads();
    if (tid < QT && my_q < q) {
      float* tk = topk[tid];
      for (int w = 0; w < RT / 32; ++w) {
        uint32_t bits = mask[tid][w];
        mask[tid][w] = 0u;
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1;
          const int rl = w * 32 + b;
          // the running threshold in row order (it fell since the tile began)
          if (!(sL[tid][rl] < tk[k - 1])) continue;
          if (emitted < static_cast<uint32_t>(cap))
            cand_idx[cand_base + emitted] = static_cast<uint32_t>(t0 + rl);
          ++emitted;
          const float u = sU[tid][rl];
          if (u < tk[k - 1]) {
            int pos = k - 1;
            while (pos > 0 && tk[pos - 1] > u) {
              tk[pos] = tk[pos - 1];
              --pos;
            }
            tk[pos] = u;
          }
        }
      }
      thr[tid] = tk[k - 1];
    }
  }
  if (tid < QT && my_q < q) {
    cand_count[static_cast<size_t>(my_q) * nseg + seg] = emitted;
    if (out_thr)
      atomicMin(reinterpret_cast<unsigned int*>(out_thr) + my_q, __float_as_uint(topk[tid][k - 1]));
  }
}

// ---------------------------------------------------------------------------
// Mixed Manhattan on 16-bit codes.  The index's values are coded as
// c = clamp(rint((v - lo) / h), 0, 65535) with h a power of two and lo a
// multiple of h, so deq(c) = lo + c h is exact in f64.  Per row the code
// error er >= sum_i |y_i - deq(c_i)| / h (an integer, rounded up); per query
// eq >= sum_i |x_i - deq(c_i)| / h (f64, rounded up; clamping included).
// With D = sum_i |x_i - y_i| and sad = sum_i |cx_i - cy_i| (v_sad_u16, two
// dimensions per instruction, exact in u32), the triangle inequality gives
//   h (sad - eq - er) <= D <= h (sad + eq + er).
// The reference sum S = f64 fold of |fl32(x_i - y_i)| (manhattan.rs:56-59)
// lies in [(1 - eps) D, (1 + eps) D], eps = 2^-24 + (d + 1) 2^-52 (one f32
// rounding per term, at most d f64 roundings of the fold), so
//   L = (1 - eps) h (sad - eq - er) <= S <= U = (1 + eps) h (sad + eq + er).
// The scan keeps per query the k smallest V = sad + er (the k smallest U)
// and stores every row with L below the running k-th U, i.e. with
//   W = sad - er < rho (V_k + eq) + eq,  rho >= (1 + eps) / (1 - eps),
// tested in integers against the right side rounded up.  Padding dimensions
// code to 0 on both sides and add nothing.
constexpr int kQ16Chunk = 8;  // dwords (16 dimensions) staged per step
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr double kQ16Limit = 1e30;

// thr = ceil(rho (vk + eq) + eq) with every operation rounded up; "no
// threshold" (fewer than k rows seen, or a query outside the domain) is
// INT_MAX, which every W is below.
__device__ __forceinline__ int q16_threshold(uint32_t vk, double eq, double rho) {
  if (vk == 0xffffffffu || !(eq < INFINITY)) return 0x7fffffff;
  const double r = __ocml_add_rtp_f64(
      __ocml_mul_rtp_f64(__ocml_add_rtp_f64(static_cast<double>(vk), eq), rho), eq);
  const double c = ceil(r);
  return c >= 2147483647.0 ? 0x7fffffff : static_cast<int>(c);
}

__global__ __launch_bounds__(NT) void knn_scan_l1q_kernel(
    const uint32_t* __restrict__ rows_q, const uint32_t* __restrict__ er, int64_t r0, int64_t n,
    int ldq, const uint32_t* __restrict__ qcodes, const double* __restrict__ eqv, int64_t q,
    int k, int64_t seg_len, int nseg, double rho, const uint32_t* __restrict__ seed,
    uint32_t* __restrict__ out_vk, uint32_t* __restrict__ cand_idx,
    uint32_t* __restrict__ cand_count, int cap) {
  extern __shared__ __align__(16) unsigned char q16_smem[];
  // queries: all their dwords, staged once; rows: kQ16Chunk dwords at a time
  // (query rows padded to 16 bytes: a lane reads its four queries' dword
  // with one 16-byte load)
  uint32_t (*Qs)[QT + 4] = reinterpret_cast<uint32_t (*)[QT + 4]>(q16_smem);
  uint32_t (*Rs)[RT + 1] = reinterpret_cast<uint32_t (*)[RT + 1]>(
      q16_smem + sizeof(uint32_t) * ldq * (QT + 4));
  __shared__ uint32_t mask[QT][RT / 32];
  __shared__ int sW[QT][RT];
  __shared__ uint32_t topk[QT][KMAX_TILED];
  __shared__ int thr[QT];

  const int tid = threadIdx.x;
  const int tq = tid >> 5;
  const int tr = tid & 31;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * QT;
  const int seg = blockIdx.y;
  const int64_t r_begin = r0 + static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);

  for (int i = tid; i < QT * KMAX_TILED; i += NT) (&topk[0][0])[i] = 0xffffffffu;
  for (int i = tid; i < QT * (RT / 32); i += NT) (&mask[0][0])[i] = 0u;
  for (int x = tid; x < QT * ldq; x += NT) {
    const int r = x / ldq, c = x % ldq;
    const int64_t row = q0 + r;
    Qs[c][r] = row < q ? qcodes[row * ldq + c] : 0u;
  }
  __syncthreads();
  const int64_t my_q = q0 + tid;
  const double my_eq = (tid < QT && my_q < q) ? eqv[my_q] : INFINITY;
  if (tid < QT) {
    // a seed is the k-th V of earlier rows: it caps every slot
    const uint32_t s0 = (seed && my_q < q) ? seed[my_q] : 0xffffffffu;
    for (int j = 0; j < k; ++j) topk[tid][j] = s0;
    thr[tid] = q16_threshold(s0, my_eq, rho);
  }
  uint32_t emitted = 0;
  const size_t cand_base = (tid < QT && my_q < q)
                               ? (static_cast<size_t>(my_q) * nseg + seg) * static_cast<size_t>(cap)
                               : 0;

  for (int64_t t0 = r_begin; t0 < r_end; t0 += RT) {
    uint32_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0u;
    for (int c0 = 0; c0 < ldq; c0 += kQ16Chunk) {
      __syncthreads();
      for (int x = tid; x < RT * kQ16Chunk; x += NT) {
        const int r = x / kQ16Chunk, c = x % kQ16Chunk;
        const int64_t row = t0 + r;
        Rs[c][r] = row < r_end ? rows_q[row * ldq + c0 + c] : 0u;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kQ16Chunk; ++kk) {
        uint32_t b[4];
        const u32x4 a = *reinterpret_cast<const u32x4*>(&Qs[c0 + kk][tq * 4]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Rs[kk][tr + 32 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_sad_u16(a[i], b[j], acc[i][j]);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rl = tr + 32 * j;
      const int64_t row = t0 + rl;
      if (row >= r_end) continue;
      const uint32_t e = er[row];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int qi = tq * 4 + i;
        const int W = static_cast<int>(acc[i][j]) - static_cast<int>(e);
        if (q0 + qi < q && W < thr[qi]) {
          sW[qi][rl] = W;
          atomicOr(&mask[qi][rl >> 5], 1u << (rl & 31));
        }
      }
    }
    __syncthreads();
    if (tid < QT && my_q < q) {
      uint32_t* tk = topk[tid];
      int th = thr[tid];
      for (int w = 0; w < RT / 32; ++w) {
        uint32_t bits = mask[tid][w];
        mask[tid][w] = 0u;
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1;
          const int rl = w * 32 + b;
          // the running threshold in row order (it fell since the tile began)
          if (!(sW[tid][rl] < th)) continue;
          if (emitted < static_cast<uint32_t>(cap))
            cand_idx[cand_base + emitted] = static_cast<uint32_t>(t0 + rl);
          ++emitted;
          const uint32_t v = static_cast<uint32_t>(sW[tid][rl] + 2 * static_cast<int>(er[t0 + rl]));
          if (v < tk[k - 1]) {
            int pos = k - 1;
            while (pos > 0 && tk[pos - 1] > v) {
              tk[pos] = tk[pos - 1];
              --pos;
            }
            tk[pos] = v;
            th = q16_threshold(tk[k - 1], my_eq, rho);
          }
        }
      }
      thr[tid] = th;
    }
  }
  if (tid < QT && my_q < q) {
    cand_count[static_cast<size_t>(my_q) * nseg + seg] = emitted;
    if (out_vk) atomicMin(out_vk + my_q, topk[tid][k - 1]);
  }
}

// Range of the index's values (first d of each row): per block min, max,
// sum |v| and a flag for values outside the form's domain.
__global__ void q16_range_kernel(const float* __restrict__ rows, int64_t n, int d, int ld,
                                 double* __restrict__ part) {
  __shared__ double smin[256], smax[256], ssum[256], sbad[256];
  double mn = INFINITY, mx = -INFINITY, sm = 0.0, bad = 0.0;
  const int64_t total = n * d;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double v = rows[(e / d) * ld + e % d];
    if (!(fabs(v) <= kQ16Limit)) {
      bad = 1.0;
      continue;
    }
    mn = fmin(mn, v);
    mx = fmax(mx, v);
    sm += fabs(v);
  }
  const int t = threadIdx.x;
  smin[t] = mn;
  smax[t] = mx;
  ssum[t] = sm;
  sbad[t] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      smin[t] = fmin(smin[t], smin[t + s]);
      smax[t] = fmax(smax[t], smax[t + s]);
      ssum[t] += ssum[t + s];
      sbad[t] = fmax(sbad[t], sbad[t + s]);
    }
    __syncthreads();
  }
  if (t == 0) {
    part[4 * blockIdx.x + 0] = smin[0];
    part[4 * blockIdx.x + 1] = smax[0];
    part[4 * blockIdx.x + 2] = ssum[0];
    part[4 * blockIdx.x + 3] = sbad[0];
  }
}

// Codes of one vector's value (f64 arithmetic; deq(c) = lo + c h exact) and
// the absolute code error.
__device__ __forceinline__ uint32_t q16_code(float v, double lo, double h, double inv_h,
                                             double* err) {
  double c = rint((static_cast<double>(v) - lo) * inv_h);
  c = fmin(fmax(c, 0.0), 65535.0);
  *err += fabs(static_cast<double>(v) - (lo + c * h));
  return static_cast<uint32_t>(c);
}

// Rows: codes (dimension pairs in dwords, padding dimensions 0) and er.
__global__ void q16_rows_kernel(const float* __restrict__ rows, int64_t n, int d, int ld,
                                double lo, double h, uint32_t* __restrict__ codes,
                                uint32_t* __restrict__ er) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rows + i * ld;
  uint32_t* o = codes + i * (ld / 2);
  const double inv_h = 1.0 / h;
  double err = 0.0;
  for (int j = 0; j < ld; j += 2) {
    const uint32_t c0 = j < d ? q16_code(r[j], lo, h, inv_h, &err) : 0u;
    const uint32_t c1 = j + 1 < d ? q16_code(r[j + 1], lo, h, inv_h, &err) : 0u;
    o[j / 2] = c0 | (c1 << 16);
  }
  // err carries at most two roundings per term (the difference, the add):
  // scale it up by more than that, then round up to an integer
  er[i] = static_cast<uint32_t>(ceil(err * inv_h * (1.0 + (2.0 * d + 4.0) * 2.220446049250313e-16)));
}

// Queries: codes (clamped) and eq; a query with a value outside the domain
// gets eq = +inf (no threshold: its lists overflow and the serial kernel
// answers it).
__global__ void q16_queries_kernel(const float* __restrict__ qs, int64_t q, int d, int ld,
                                   double lo, double h, uint32_t* __restrict__ codes,
                                   double* __restrict__ eq) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= q) return;
  const float* r = qs + i * ld;
  uint32_t* o = codes + i * (ld / 2);
  const double inv_h = 1.0 / h;
  double err = 0.0;
  bool ok = true;
  for (int j = 0; j < d; ++j) ok = ok && fabsf(r[j]) <= static_cast<float>(kQ16Limit);
  for (int j = 0; j < ld; j += 2) {
    const uint32_t c0 = j < d ? q16_code(r[j], lo, h, inv_h, &err) : 0u;
    const uint32_t c1 = j + 1 < d ? q16_code(r[j + 1], lo, h, inv_h, &err) : 0u;
    o[j / 2] = c0 | (c1 << 16);
  }
  eq[i] = ok ? __ocml_mul_rtp_f64(err * inv_h, 1.0 + (2.0 * d + 4.0) * 2.220446049250313e-16)
             : INFINITY;
}

// The reference sum of every stored candidate (one workgroup per (query,
// segment) list, one thread per candidate).
template <typename T, int M>
__global__ void knn_rescore_kernel(const T* __restrict__ refs, int ld, const T* __restrict__ queries,
                                   int64_t q, double p, int nseg, int cap,
                                   const uint32_t* __restrict__ cand_idx,
                                   const uint32_t* __restrict__ cand_count,
                                   double* __restrict__ cand_s) {
  const int64_t list = blockIdx.x;
  const int64_t qi = list / nseg;
  if (qi >= q) return;
  const uint32_t cnt = min(cand_count[list], static_cast<uint32_t>(cap));
  const T* qrow = queries + qi * ld;
  const size_t base = static_cast<size_t>(list) * static_cast<size_t>(cap);
  for (uint32_t c = threadIdx.x; c < cnt; c += blockDim.x) {
    const T* rrow = refs + static_cast<int64_t>(cand_idx[base + c]) * ld;
    double s = 0.0;
    for (int j = 0; j < ld; ++j) s = s + Term<T, M>::f(qrow[j], rrow[j], p);
    cand_s[base + c] = s;
  }
}

// A candidate list set: per query `nseg` lists of up to `cap` (row, s)
// entries, each list in row order and the lists in row order.
struct CandSet {
  const uint32_t* idx;
  const double* s;
  const uint32_t* count;
  int nseg;
  int cap;
};

// HeapSelection's replay over the stored candidates, one wavefront per
// query: the earlier rows' lists (an optional prefix set), then the main
// set, each in row order.  Entries are taken 64 at a time (lane = entry);
// those whose distance is below the heap's largest (the reference's strict
// <) enter one after another, in row order, by a sift-down of the heap
// held across the lanes, and the rest are tested again against the new top.
constexpr int kReplayWaves = 4;
__device__ __forceinline__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(__double_as_longlong(v)), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(__double_as_longlong(v) >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned long long>(static_cast<uint32_t>(lo)));
}

// The reference heap held across a wave's lanes: lane j < k holds entry j
// (distance hd, row hi); root replacement and HeapSelection::sift_down
// (heap_select.rs:71-86) with readlane reads and select swaps.
__device__ __forceinline__ void lane_heap_push_root(double& hd, uint32_t& hi, int lane, int k,
                                                    double d, uint32_t row) {
  if (lane == 0) {
    hd = d;
    hi = row;
  }
  const int n = k - 1;
  int kk = 0;
  const double vk = d;
  while (2 * kk <= n) {
    int j = 2 * kk;
    double vj = lane_value(hd, j);
    if (j < n) {
      const double vj1 = lane_value(hd, j + 1);
      if (vj < vj1) {
        j += 1;
        vj = vj1;
      }
    }
    if (vk == vj || vk > vj) break;
    const uint32_t ij = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(hi), j));
    const uint32_t ik = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(hi), kk));
    if (lane == kk) {
      hd = vj;
      hi = ij;
    } else if (lane == j) {
      hd = vk;
      hi = ik;
    }
    kk = j;
  }
}

template <int M>
__global__ __launch_bounds__(64 * kReplayWaves) void knn_replay_kernel(
    int64_t q, int k, double p, CandSet pre, int nseg, int cap,
    const uint32_t* __restrict__ cand_idx, const double* __restrict__ cand_s,
    const uint32_t* __restrict__ cand_count, int64_t* __restrict__ out_idx,
    double* __restrict__ out_dist, uint32_t* __restrict__ out_count,
    uint32_t* __restrict__ overflow_list, uint32_t* __restrict__ overflow_n,
    unsigned long long* __restrict__ total_cand) {
  const int lane = threadIdx.x & 63;
  const int64_t qi = static_cast<int64_t>(blockIdx.x) * kReplayWaves + (threadIdx.x >> 6);
  if (qi >= q) return;  // the whole wave
  bool over = false;
  for (int s2 = lane; s2 < pre.nseg; s2 += 64)
    over = over || pre.count[static_cast<size_t>(qi) * pre.nseg + s2] > static_cast<uint32_t>(pre.cap);
  for (int s2 = lane; s2 < nseg; s2 += 64)
    over = over || cand_count[static_cast<size_t>(qi) * nseg + s2] > static_cast<uint32_t>(cap);
  if (__ballot(over) != 0) {
    if (lane == 0) {
      const uint32_t slot = atomicAdd(overflow_n, 1u);
      overflow_list[slot] = static_cast<uint32_t>(qi);
      out_count[qi] = 0;
    }
    return;
  }
  // the heap in lanes 0 .. k-1 (seeded with k x {inf, None})
  double hd = INFINITY;
  uint32_t hi = 0xffffffffu;
  double top = INFINITY;
  unsigned long long seen = 0;
  auto walk = [&](const uint32_t* idx, const double* sv, size_t base, uint32_t cnt) {
    seen += cnt;
    for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {
      const bool valid = c0 + lane < cnt;
      const double d = valid ? finalize<M>(sv[base + c0 + lane], p) : 0.0;
      const uint32_t row = valid ? idx[base + c0 + lane] : 0u;
      uint64_t m = __ballot(valid && d < top);
      while (m) {
        const int src = __builtin_ctzll(m);
        const double dl = lane_value(d, src);
        const uint32_t rl = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(row), src));
        lane_heap_push_root(hd, hi, lane, k, dl, rl);
        top = lane_value(hd, 0);
        m &= m - 1;
        m &= __ballot(valid && d < top);
      }
    }
  };
  for (int s2 = 0; s2 < pre.nseg; ++s2) {
    const size_t base = (static_cast<size_t>(qi) * pre.nseg + s2) * static_cast<size_t>(pre.cap);
    walk(pre.idx, pre.s, base, pre.count[static_cast<size_t>(qi) * pre.nseg + s2]);
  }
  for (int s2 = 0; s2 < nseg; ++s2) {
    const size_t base = (static_cast<size_t>(qi) * nseg + s2) * static_cast<size_t>(cap);
    walk(cand_idx, cand_s, base, cand_count[static_cast<size_t>(qi) * nseg + s2]);
  }
  if (lane == 0) atomicAdd(total_cand, seen);
  // linear_search.rs:79-83: keep heap-array order, drop None entries.
  const bool full = lane < k && hi != 0xffffffffu;
  const uint64_t fm = __ballot(full);
  const int c = __popcll(fm);
  if (full) {
    const int pos = __popcll(fm & ((uint64_t(1) << lane) - 1));
    out_idx[qi * k + pos] = static_cast<int64_t>(hi);
    out_dist[qi * k + pos] = hd;
  }
  if (lane >= c && lane < k) {
    out_idx[qi * k + lane] = -1;
    out_dist[qi * k + lane] = NAN;
  }
  if (lane == 0) out_count[qi] = static_cast<uint32_t>(c);
}

// One workgroup per listed query; persistent over the list so the grid can
// be launched before the list length is known.
template <typename T, int M>
__global__ __launch_bounds__(NT) void knn_serial_kernel(
    const T* __restrict__ refs, int64_t n, int ld, const T* __restrict__ queries, int k, double p,
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_n, int64_t* __restrict__ out_idx,
    double* __restrict__ out_dist, uint32_t* __restrict__ out_count) {
  __shared__ double dvals[NT];
  __shared__ uint32_t flags[NT / 32];
  __shared__ double root;
  const int tid = threadIdx.x;
  const uint32_t total = *list_n;
  for (uint32_t li = blockIdx.x; li < total; li += gridDim.x) {
    const int64_t qi = list[li];
    const T* qrow = queries + qi * ld;
    double* hd = out_dist + qi * k;
    int64_t* hi = out_idx + qi * k;
    if (tid == 0) {
      for (int j = 0; j < k; ++j) {
        hd[j] = INFINITY;
        hi[j] = -1;
      }
      root = INFINITY;
    }
    if (tid < NT / 32) flags[tid] = 0u;
    __syncthreads();
    for (int64_t c0 = 0; c0 < n; c0 += NT) {
      const int64_t r = c0 + tid;
      double d = NAN;
      if (r < n) {
        const T* rrow = refs + r * ld;
        double s = 0.0;
        for (int j = 0; j < ld; ++j) s = s + Term<T, M>::f(qrow[j], rrow[j], p);
        d = finalize<M>(s, p);
      }
      const double cur = root;
      dvals[tid] = d;
      if (d < cur) atomicOr(&flags[tid >> 5], 1u << (tid & 31));
      __syncthreads();
      if (tid == 0) {
        for (int w = 0; w < NT / 32; ++w) {
          uint32_t bits = flags[w];
          flags[w] = 0u;
          while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            const int t = w * 32 + b;
            const double dv = dvals[t];
            if (dv < hd[0]) heap_push_root(hd, hi, k, dv, c0 + t);
          }
        }
        root = hd[0];
      }
      __syncthreads();
    }
    if (tid == 0) {
      int c = 0;
      for (int j = 0; j < k; ++j) {
        if (hi[j] >= 0) {
          hi[c] = hi[j];
          hd[c] = hd[j];
          ++c;
        }
      }
      for (int j = c; j < k; ++j) {
        hi[j] = -1;
        hd[j] = NAN;
      }
      out_count[qi] = static_cast<uint32_t>(c);
    }
    __syncthreads();
  }
}

__global__ void iota_list_kernel(uint32_t* list, uint32_t* list_n, int64_t q) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < q) list[i] = static_cast<uint32_t>(i);
  if (i == 0) *list_n = static_cast<uint32_t>(q);
}

struct Plan {
  int nseg;
  int64_t seg_len;
  int cap;
};

Plan make_plan(int64_t n, int64_t q, int k) {
  const int64_t nqt = (q + QT - 1) / QT;
  const int64_t tiles = (n + RT - 1) / RT;
  int64_t nseg = std::max<int64_t>(1, (4096 + nqt - 1) / nqt);
  nseg = std::min<int64_t>(nseg, tiles);
  nseg = std::min<int64_t>(nseg, 65535);
  int64_t seg_tiles = (tiles + nseg - 1) / nseg;
  nseg = (tiles + seg_tiles - 1) / seg_tiles;
  const int64_t seg_len = seg_tiles * RT;
  // Expected number of running-k-th improvements over L rows is about
  // k (1 + ln(L / k)); keep three times that plus slack.
  const double L = static_cast<double>(std::min<int64_t>(seg_len, n));
  double expect = k * (1.0 + std::log(std::max(1.0, L / k)));
  int64_t cap = static_cast<int64_t>(3.0 * expect) + 64;
  cap = std::min<int64_t>(cap, seg_len);
  return Plan{static_cast<int>(nseg), seg_len, static_cast<int>(cap)};
}

template <typename T, int M>
int run_exact(const sc_index* index, const T* d_queries, int64_t q, int k, double p, int64_t* d_idx,
              double* d_dist, uint32_t* d_count, hipStream_t stream, sc_stats* stats) {
  const int64_t n = static_cast<int64_t>(index->n);
  const int ld = static_cast<int>(index->ld);
  const T* refs = static_cast<const T*>(index->rows);

  DevBuf flags;  // [0] overflow count, [1..2] candidate total (u64)
  SC_HIP(flags.alloc(64, stream));
  SC_HIP(hipMemsetAsync(flags.p, 0, 64, stream));
  uint32_t* overflow_n = flags.as<uint32_t>();
  unsigned long long* total_cand = reinterpret_cast<unsigned long long*>(flags.as<char>() + 8);
  DevBuf olist;
  SC_HIP(olist.alloc(sizeof(uint32_t) * std::max<int64_t>(q, 1), stream));

  if (k <= KMAX_TILED) {
    // Bound the candidate store to ~1 GiB by processing query batches.
    Plan plan = make_plan(n, q, k);
    const size_t per_query = static_cast<size_t>(plan.nseg) * plan.cap * 12 + plan.nseg * 4;
    int64_t qb = std::max<int64_t>(QT, static_cast<int64_t>((size_t(1) << 30) / per_query) / QT * QT);
    qb = std::min(qb, q);
    DevBuf cidx, cs, ccnt;
    SC_HIP(cidx.alloc(sizeof(uint32_t) * qb * plan.nseg * plan.cap, stream));
    SC_HIP(cs.alloc(sizeof(double) * qb * plan.nseg * plan.cap, stream));
    SC_HIP(ccnt.alloc(sizeof(uint32_t) * qb * plan.nseg, stream));
    for (int64_t b0 = 0; b0 < q; b0 += qb) {
      const int64_t qn = std::min(qb, q - b0);
      dim3 grid(static_cast<unsigned>((qn + QT - 1) / QT), static_cast<unsigned>(plan.nseg));
      {
        ProfScope ps("knn_scan", stream);
        hipLaunchKernelGGL((knn_scan_kernel<T, M>), grid, dim3(NT), 0, stream, refs, n, ld,
                           static_cast<int>(index->d), d_queries + b0 * ld, qn, k, p,
                           plan.seg_len, plan.nseg,
                           cidx.as<uint32_t>(), cs.as<double>(), ccnt.as<uint32_t>(), plan.cap);
        SC_HIP(hipGetLastError());
      }
      {
        ProfScope ps("knn_replay", stream);
        hipLaunchKernelGGL((knn_replay_kernel<M>),
                           dim3(static_cast<unsigned>((qn + kReplayWaves - 1) / kReplayWaves)),
                           dim3(64 * kReplayWaves), 0, stream, qn, k, p, CandSet{nullptr, nullptr, nullptr, 0, 1},
                           plan.nseg, plan.cap,
                           cidx.as<uint32_t>(), cs.as<double>(), ccnt.as<uint32_t>(),
                           d_idx + b0 * k, d_dist + b0 * k, d_count + b0, olist.as<uint32_t>(),
                           overflow_n, total_cand);
        SC_HIP(hipGetLastError());
      }
      {
        // batch-relative query ids in the overflow list
        ProfScope ps("knn_serial", stream);
        hipLaunchKernelGGL((knn_serial_kernel<T, M>), dim3(512), dim3(NT), 0, stream, refs, n,
                           ld, d_queries + b0 * ld, k, p, olist.as<uint32_t>(), overflow_n,
                           d_idx + b0 * k, d_dist + b0 * k, d_count + b0);
        SC_HIP(hipGetLastError());
      }
      if (stats) {
        uint32_t ov = 0;
        SC_HIP(hipMemcpyAsync(&ov, overflow_n, 4, hipMemcpyDeviceToHost, stream));
        SC_HIP(hipStreamSynchronize(stream));
        stats->overflow_queries += ov;
      }
      SC_HIP(hipMemsetAsync(overflow_n, 0, 4, stream));
    }
  } else {
    hipLaunchKernelGGL(iota_list_kernel, dim3(static_cast<unsigned>((q + 255) / 256)), dim3(256), 0,
                       stream, olist.as<uint32_t>(), overflow_n, q);
    SC_HIP(hipGetLastError());
    hipLaunchKernelGGL((knn_serial_kernel<T, M>), dim3(1024), dim3(NT), 0, stream, refs, n, ld,
                       d_queries, k, p, olist.as<uint32_t>(), overflow_n, d_idx, d_dist, d_count);
    SC_HIP(hipGetLastError());
    if (stats) stats->overflow_queries += static_cast<uint64_t>(q);
  }
  if (stats) {
    unsigned long long tc = 0;
    SC_HIP(hipMemcpyAsync(&tc, total_cand, 8, hipMemcpyDeviceToHost, stream));
    SC_HIP(hipStreamSynchronize(stream));
    stats->candidates += tc;
    stats->queries += static_cast<uint64_t>(q);
  }
  return SC_OK;
}

// SC_MIXED Manhattan on f32 rows: bounded f32 scan, rescoring in the
// reference arithmetic, replay; overflowing queries take the serial kernel.
// Indexes of at least 4 x 65,536 rows scan their first 65,536 rows first, in
// one segment or (for few queries) up to 64; the smallest of the segments'
// k-th U seeds the threshold of every later segment (each is at or above
// the k-th S of the whole prefix, hence at or above the heap's threshold
// from that row on), so the later segments store only the rows that can
// still enter.
constexpr int64_t kL1Prefix = 65536;

// Builds the 16-bit coded copy of an f32 index once (under its lock).  The
// form is taken when every value is finite with |v| <= 1e30 and the code step
// h (the smallest power of two with (max - min) / h <= 65534) is at most
// 1/64 of the mean |v|, so the bound's band (about d h) stays a few percent
// of a typical distance; otherwise, and for rows so wide that the staged
// queries would not fit the LDS, the index keeps the f32 scan.
constexpr int kQ16MaxLdq = 416;  // dwords of a staged query row (LDS: 36 x 416 x 4 B)

int ensure_q16(sc_index* ix, hipStream_t s) {
  std::lock_guard<std::mutex> lk(*static_cast<std::mutex*>(ix->mu));
  if (ix->q16_state != 0) return SC_OK;
  const int64_t n = static_cast<int64_t>(ix->n);
  const int d = static_cast<int>(ix->d), ld = static_cast<int>(ix->ld);
  if (n == 0 || d == 0 || ld / 2 > kQ16MaxLdq) {
    ix->q16_state = 2;
    return SC_OK;
  }
  const float* rows = static_cast<const float*>(ix->rows);
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (n * d + 255) / 256));
  DevBuf part;
  SC_HIP(part.alloc(sizeof(double) * 4 * blocks, s));
  hipLaunchKernelGGL(q16_range_kernel, dim3(blocks), dim3(256), 0, s, rows, n, d, ld,
                     part.as<double>());
  SC_HIP(hipGetLastError());
  std::vector<double> h(4 * static_cast<size_t>(blocks));
  SC_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  double mn = INFINITY, mx = -INFINITY, sum = 0.0, bad = 0.0;
  for (int b = 0; b < blocks; ++b) {
    mn = std::min(mn, h[4 * b]);
    mx = std::max(mx, h[4 * b + 1]);
    sum += h[4 * b + 2];
    bad = std::max(bad, h[4 * b + 3]);
  }
  if (bad != 0.0 || !(mn <= mx)) {
    ix->q16_state = 2;
    return SC_OK;
  }
  const double range = mx - mn;
  double step = std::ldexp(1.0, -100);
  if (range > 0.0) {
    int e = 0;
    (void)std::frexp(range / 65534.0, &e);  // range / 65534 < 2^e
    step = std::max(step, std::ldexp(1.0, e));
  }
  double lo = std::floor(mn / step) * step;
  while ((mx - lo) / step > 65535.0) {  // not reached for the step above
    step *= 2.0;
    lo = std::floor(mn / step) * step;
  }
  const double mean_abs = sum / (static_cast<double>(n) * d);
  if (range > 0.0 && step > mean_abs / 64.0) {
    ix->q16_state = 2;
    return SC_OK;
  }
  void* codes = nullptr;
  uint32_t* er = nullptr;
  SC_HIP(hipMalloc(&codes, sizeof(uint32_t) * static_cast<size_t>(n) * (ld / 2)));
  if (hipMalloc(reinterpret_cast<void**>(&er), sizeof(uint32_t) * n) != hipSuccess) {
    (void)hipFree(codes);
    return hip_fail(hipErrorOutOfMemory, "hipMalloc(row code errors)");
  }
  hipLaunchKernelGGL(q16_rows_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                     rows, n, d, ld, lo, step, static_cast<uint32_t*>(codes), er);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(codes);
    (void)hipFree(er);
    return hip_fail(e, "q16_rows_kernel");
  }
  ix->rows_q16 = codes;
  ix->err_q16 = er;
  ix->q16_lo = lo;
  ix->q16_step = step;
  ix->q16_state = 1;
  return SC_OK;
}

int run_l1_mixed(const sc_index* index, const float* d_queries, int64_t q, int k, int64_t* d_idx,
                 double* d_dist, uint32_t* d_count, hipStream_t stream, sc_stats* stats) {
  if (k > KMAX_TILED)
    return run_exact<float, kManhattan>(index, d_queries, q, k, 1.0, d_idx, d_dist, d_count,
                                        stream, stats);
  const int64_t n = static_cast<int64_t>(index->n);
  const int ld = static_cast<int>(index->ld);
  const float* refs = static_cast<const float*>(index->rows);
  if (int rc = ensure_q16(const_cast<sc_index*>(index), stream)) return rc;
  const bool coded = index->q16_state == 1;
  const int ldq = ld / 2;
  // rho >= (1 + eps) / (1 - eps), eps = 2^-24 + (d + 1) 2^-52
  const double eps = 5.9604644775390625e-08 + (static_cast<double>(index->d) + 1.0) *
                                                  2.220446049250313e-16;
  const double rho = 1.0 + 3.0 * eps;
  const size_t q16_lds = sizeof(uint32_t) * (static_cast<size_t>(ldq) * (QT + 4) +
                                             static_cast<size_t>(kQ16Chunk) * (RT + 1));
  DevBuf flags;
  SC_HIP(flags.alloc(64, stream));
  SC_HIP(hipMemsetAsync(flags.p, 0, 64, stream));
  uint32_t* overflow_n = flags.as<uint32_t>();
  unsigned long long* total_cand = reinterpret_cast<unsigned long long*>(flags.as<char>() + 8);
  DevBuf olist;
  SC_HIP(olist.alloc(sizeof(uint32_t) * std::max<int64_t>(q, 1), stream));
  const int64_t prefix = n >= 4 * kL1Prefix ? kL1Prefix : 0;
  // The prefix runs in nseg_p segments (more when few query tiles would
  // leave the chip idle); each segment's k-th bound caps the k-th sum of all
  // the prefix rows, so the smallest of them seeds every later segment.
  const int64_t qtiles_all = (q + QT - 1) / QT;
  int nseg_p = 1;  // a power of two, so the segments tile the prefix exactly
  while (prefix && nseg_p < 64 && qtiles_all * nseg_p * 2 <= 256) nseg_p *= 2;
  const int64_t seg_p = prefix / nseg_p;
  const int cap_p = prefix ? static_cast<int>(std::min<int64_t>(
                                 seg_p, static_cast<int64_t>(3.0 * k * (1.0 + std::log(std::max(
                                            1.0, static_cast<double>(seg_p) / k)))) + 64))
                           : 1;
  Plan plan = make_plan(n - prefix, q, k);
  const size_t per_query = static_cast<size_t>(plan.nseg) * plan.cap * 12 + plan.nseg * 4 +
                           (prefix ? static_cast<size_t>(nseg_p) * (cap_p * 12 + 4) + 4 : 0);
  int64_t qb = std::max<int64_t>(QT, static_cast<int64_t>((size_t(1) << 30) / per_query) / QT * QT);
  qb = std::min(qb, q);
  DevBuf cidx, cs, ccnt, pidx, ps, pcnt, seed, qcodes, qerr;
  if (coded) {
    SC_HIP(qcodes.alloc(sizeof(uint32_t) * qb * ldq, stream));
    SC_HIP(qerr.alloc(sizeof(double) * qb, stream));
  }
  SC_HIP(cidx.alloc(sizeof(uint32_t) * qb * plan.nseg * plan.cap, stream));
  SC_HIP(cs.alloc(sizeof(double) * qb * plan.nseg * plan.cap, stream));
  SC_HIP(ccnt.alloc(sizeof(uint32_t) * qb * plan.nseg, stream));
  if (prefix) {
    SC_HIP(pidx.alloc(sizeof(uint32_t) * qb * nseg_p * cap_p, stream));
    SC_HIP(ps.alloc(sizeof(double) * qb * nseg_p * cap_p, stream));
    SC_HIP(pcnt.alloc(sizeof(uint32_t) * qb * nseg_p, stream));
    SC_HIP(seed.alloc(sizeof(float) * qb, stream));
  }
  for (int64_t b0 = 0; b0 < q; b0 += qb) {
    const int64_t qn = std::min(qb, q - b0);
    const float* qs = d_queries + b0 * ld;
    const unsigned qtiles = static_cast<unsigned>((qn + QT - 1) / QT);
    if (coded) {
      hipLaunchKernelGGL(q16_queries_kernel, dim3(static_cast<unsigned>((qn + 255) / 256)),
                         dim3(256), 0, stream, qs, qn, static_cast<int>(index->d), ld,
                         index->q16_lo, index->q16_step, qcodes.as<uint32_t>(), qerr.as<double>());
      SC_HIP(hipGetLastError());
      ProfScope pscope("knn_scan_q16", stream);
      const uint32_t* rq = static_cast<const uint32_t*>(index->rows_q16);
      if (prefix) {
        // the segments' k-th V meet by atomicMin: start at "none"
        SC_HIP(hipMemsetAsync(seed.p, 0xff, sizeof(uint32_t) * qn, stream));
        hipLaunchKernelGGL(knn_scan_l1q_kernel, dim3(qtiles, static_cast<unsigned>(nseg_p)),
                           dim3(NT), q16_lds, stream, rq, index->err_q16, int64_t{0}, prefix, ldq,
                           qcodes.as<uint32_t>(), qerr.as<double>(), qn, k, seg_p, nseg_p, rho,
                           static_cast<const uint32_t*>(nullptr), seed.as<uint32_t>(),
                           pidx.as<uint32_t>(), pcnt.as<uint32_t>(), cap_p);
        SC_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(knn_scan_l1q_kernel, dim3(qtiles, static_cast<unsigned>(plan.nseg)),
                         dim3(NT), q16_lds, stream, rq, index->err_q16, prefix, n, ldq,
                         qcodes.as<uint32_t>(), qerr.as<double>(), qn, k, plan.seg_len, plan.nseg,
                         rho, prefix ? static_cast<const uint32_t*>(seed.as<uint32_t>()) : nullptr,
                         static_cast<uint32_t*>(nullptr), cidx.as<uint32_t>(), ccnt.as<uint32_t>(),
                         plan.cap);
      SC_HIP(hipGetLastError());
    } else {
      ProfScope pscope("knn_scan", stream);
      if (prefix) {
        // the segments' k-th U (>= 0) meet by atomicMin on their bits
        SC_HIP(hipMemsetAsync(seed.p, 0xff, sizeof(float) * qn, stream));
        hipLaunchKernelGGL(knn_scan_l1f32_kernel, dim3(qtiles, static_cast<unsigned>(nseg_p)),
                           dim3(NT), 0, stream, refs, int64_t{0}, prefix, ld, qs, qn, k, seg_p,
                           nseg_p, static_cast<const float*>(nullptr), seed.as<float>(),
                           pidx.as<uint32_t>(), pcnt.as<uint32_t>(), cap_p);
        SC_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(knn_scan_l1f32_kernel, dim3(qtiles, static_cast<unsigned>(plan.nseg)),
                         dim3(NT), 0, stream, refs, prefix, n, ld, qs, qn, k, plan.seg_len,
                         plan.nseg, prefix ? static_cast<const float*>(seed.as<float>()) : nullptr,
                         static_cast<float*>(nullptr), cidx.as<uint32_t>(), ccnt.as<uint32_t>(),
                         plan.cap);
      SC_HIP(hipGetLastError());
    }
    {
      ProfScope pscope("knn_mixed_rescore", stream);
      if (prefix) {
        hipLaunchKernelGGL((knn_rescore_kernel<float, kManhattan>),
                           dim3(static_cast<unsigned>(qn * nseg_p)), dim3(64), 0, stream, refs, ld,
                           qs, qn, 1.0, nseg_p, cap_p, pidx.as<uint32_t>(), pcnt.as<uint32_t>(),
                           ps.as<double>());
        SC_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL((knn_rescore_kernel<float, kManhattan>),
                         dim3(static_cast<unsigned>(qn * plan.nseg)), dim3(64), 0, stream, refs, ld,
                         qs, qn, 1.0, plan.nseg, plan.cap, cidx.as<uint32_t>(),
                         ccnt.as<uint32_t>(), cs.as<double>());
      SC_HIP(hipGetLastError());
    }
    {
      ProfScope pscope("knn_replay", stream);
      const CandSet pre = prefix ? CandSet{pidx.as<uint32_t>(), ps.as<double>(), pcnt.as<uint32_t>(),
                                           nseg_p, cap_p}
                                 : CandSet{nullptr, nullptr, nullptr, 0, 1};
      hipLaunchKernelGGL((knn_replay_kernel<kManhattan>),
                         dim3(static_cast<unsigned>((qn + kReplayWaves - 1) / kReplayWaves)),
                         dim3(64 * kReplayWaves), 0, stream, qn, k, 1.0, pre, plan.nseg, plan.cap,
                         cidx.as<uint32_t>(), cs.as<double>(), ccnt.as<uint32_t>(),
                         d_idx + b0 * k, d_dist + b0 * k, d_count + b0, olist.as<uint32_t>(),
                         overflow_n, total_cand);
      SC_HIP(hipGetLastError());
    }
    {
      ProfScope pscope("knn_serial", stream);
      hipLaunchKernelGGL((knn_serial_kernel<float, kManhattan>), dim3(512), dim3(NT), 0, stream,
                         refs, n, ld, qs, k, 1.0, olist.as<uint32_t>(), overflow_n, d_idx + b0 * k,
                         d_dist + b0 * k, d_count + b0);
      SC_HIP(hipGetLastError());
    }
    if (stats) {
      uint32_t ov = 0;
      SC_HIP(hipMemcpyAsync(&ov, overflow_n, 4, hipMemcpyDeviceToHost, stream));
      SC_HIP(hipStreamSynchronize(stream));
      stats->overflow_queries += ov;
    }
    SC_HIP(hipMemsetAsync(overflow_n, 0, 4, stream));
  }
  if (stats) {
    unsigned long long tc = 0;
    SC_HIP(hipMemcpyAsync(&tc, total_cand, 8, hipMemcpyDeviceToHost, stream));
    SC_HIP(hipStreamSynchronize(stream));
    stats->candidates += tc;
    stats->shortlist_rescored += tc;
    stats->queries += static_cast<uint64_t>(q);
  }
  return SC_OK;
}

template <typename T>
int dispatch_metric(const sc_index* index, const void* d_queries, size_t q, size_t k,
                    sc_metric metric, int p, int64_t* d_idx, double* d_dist, uint32_t* d_count,
                    hipStream_t stream, sc_stats* stats) {
  const T* qs = static_cast<const T*>(d_queries);
  switch (metric) {
    case SC_EUCLIDIAN:
      return run_exact<T, kEuclid>(index, qs, q, static_cast<int>(k), 2.0, d_idx, d_dist, d_count,
                                   stream, stats);
    case SC_MANHATTAN:
      return run_exact<T, kManhattan>(index, qs, q, static_cast<int>(k), 1.0, d_idx, d_dist,
                                      d_count, stream, stats);
    case SC_MINKOWSKI:
      return run_exact<T, kMinkowski>(index, qs, q, static_cast<int>(k), static_cast<double>(p),
                                      d_idx, d_dist, d_count, stream, stats);
  }
  return fail(SC_ERR_UNSUPPORTED, "unknown metric");
}

}  // namespace

int launch_serial_exact_f32(const sc_index* index, const float* d_queries, int k,
                            const uint32_t* list, const uint32_t* list_n, int64_t* d_idx,
                            double* d_dist, uint32_t* d_count, hipStream_t stream) {
  ProfScope ps("knn_serial", stream);
  hipLaunchKernelGGL((knn_serial_kernel<float, kEuclid>), dim3(512), dim3(NT), 0, stream,
                     static_cast<const float*>(index->rows), static_cast<int64_t>(index->n),
                     static_cast<int>(index->ld), d_queries, k, 2.0, list, list_n, d_idx, d_dist,
                     d_count);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int knn_l1_mixed_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                        int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t stream,
                        sc_stats* stats) {
  if (q == 0) return SC_OK;
  if (index->dtype != SC_F32)
    return dispatch_metric<double>(index, d_queries, q, k, SC_MANHATTAN, 1, d_idx, d_dist, d_count,
                                   stream, stats);
  return run_l1_mixed(index, static_cast<const float*>(d_queries), static_cast<int64_t>(q),
                      static_cast<int>(k), d_idx, d_dist, d_count, stream, stats);
}

int knn_exact_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                     sc_metric metric, int p, int64_t* d_idx, double* d_dist, uint32_t* d_count,
                     hipStream_t stream, sc_stats* stats) {
  if (q == 0) return SC_OK;
  if (index->dtype == SC_F64)
    return dispatch_metric<double>(index, d_queries, q, k, metric, p, d_idx, d_dist, d_count,
                                   stream, stats);
  return dispatch_metric<float>(index, d_queries, q, k, metric, p, d_idx, d_dist, d_count, stream,
                                stats);
}

}  // namespace sc
