// Certified BF16-MFMA Lloyd step for f32 rows (sc_kmeans_run, SC_MIXED).
//
// One fused pass over the rows per iteration:
//   * a workgroup owns a block of kBlockRows rows and stages 128 rows at a
//     time through LDS (one HBM read of the f32 rows per iteration);
//   * each wave converts its 32 rows to BF16 B-fragments in registers and
//     multiplies them with the BF16 centroids (A-fragments, loaded once per
//     iteration) on v_mfma_f32_32x32x16_bf16, giving approximate
//     a_c = ||c||^2 - 2 x.c for every centroid;
//   * the label is certified: with E bounding |(s_c - ||x||^2) - a_c| for every
//     c (s_c the f64 squared distance KMeans::predict computes,
//     kmeans.rs:338-346), a row whose only centroid with a_c <= min a + 2E is
//     c* has s_c > s_{c*} for every other c, so c* is the exact argmin.  Rows
//     with more than one such centroid (or non-finite values) are assigned in
//     FP64 by the wave, lanes = centroids, the same arithmetic and
//     lowest-index tie rule as the FP64 assign kernel;
//   * the block's per-cluster f64 sums are accumulated exactly as the FP64
//     form's partial kernel does (owner wave per cluster, rows in order), so
//     both forms produce bit-identical partials, sums, centroids and
//     distortions whenever their labels agree -- which certification
//     guarantees.
//
// Bound: x.c - C <= c1 ||x|| ||c||, c1 = 2^-7 + 2^-16 + 2^-20 + 128 * 2^-22
// (BF16 rounding of both operands -- the centroid's through f32 --, FP32
// accumulation at four ulps per step), and every other rounding (f32 norms,
// the epilogue FMA, the f64 sum behind s_c) below 2^-20 (||x||^2 + max ||c||^2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "kmeans.hpp"
#include "metric.hpp"

namespace sc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int kMaxCB = 4;  // centroid blocks of 32 -> k <= 128
constexpr int kMaxKS = 8;  // K steps of 16 -> dimension <= 128
constexpr double kC1 = 0.0078125 + 1.52587890625e-05 + 9.5367431640625e-07 + 128.0 * 2.384185791015625e-07;
constexpr double kKappa = 9.5367431640625e-07;  // 2^-20
constexpr uint32_t kUncert = 0xffffffffu;

__device__ __forceinline__ uint32_t bf16_bits_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ void lexmin_f(float& v, uint32_t& i, float v2, uint32_t i2) {
  if (v2 < v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__device__ __forceinline__ void lexmin_d(double& v, uint32_t& i, double v2, uint32_t i2) {
  if (v2 < v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

// Centroid terms for one iteration: BF16 copy (kpad x ld), c2 = RN(||c||^2)
// (+inf for padding rows), and two scalars: RU(max ||c||), RU(max ||c||^2).
__device__ __forceinline__ void km_prep_centroids_body(const double* __restrict__ cent, int k,
                                                       int kpad, int ld,
                                                       uint16_t* __restrict__ cb16,
                                                       float* __restrict__ c2,
                                                       float* __restrict__ cmax, double* red) {
  const int tid = threadIdx.x;
  double mx = 0.0;
  for (int c = tid; c < kpad; c += NT) {
    double s = 0.0;
    for (int j = 0; j < ld; ++j) {
      const double v = c < k ? cent[static_cast<int64_t>(c) * ld + j] : 0.0;
      s += v * v;
      cb16[static_cast<int64_t>(c) * ld + j] =
          static_cast<uint16_t>(bf16_bits_rne(static_cast<float>(v)));
    }
    c2[c] = c < k ? static_cast<float>(s) : INFINITY;
    if (c < k && s < 1e300) mx = fmax(mx, s);  // NaN centroids never count
  }
  red[tid] = mx;
  __syncthreads();
  for (int w = NT / 2; w >= 1; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) {
    cmax[0] = __double2float_ru(sqrt(red[0]) * (1.0 + 9.5367431640625e-07));
    cmax[1] = __double2float_ru(red[0] * (1.0 + 9.5367431640625e-07));
  }
  __syncthreads();
}

__global__ __launch_bounds__(NT) void km_prep_centroids_kernel(const double* __restrict__ cent,
                                                               int k, int kpad, int ld,
                                                               uint16_t* __restrict__ cb16,
                                                               float* __restrict__ c2,
                                                               float* __restrict__ cmax,
                                                               const KmState* __restrict__ st) {
  if (st->done) return;
  __shared__ double red[NT];
  km_prep_centroids_body(cent, k, kpad, ld, cb16, c2, cmax, red);
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// Dynamic LDS: acc[k*d] f64 | stage[kChunkRows*ld] f32 | lab[kChunkRows] | cnt[k].
// One block b (kBlockRows rows) of one problem: certified labels and the
// block's per-cluster partial sums.  Ends with a workgroup barrier so LDS
// can be reused for the next block.
__device__ __forceinline__ void km_block_step(
    int64_t b, unsigned char* smem, const float* __restrict__ rows, int64_t n, int ld, int d,
    const double* __restrict__ cent, const uint16_t* __restrict__ cb16,
    const float* __restrict__ c2g, const float* __restrict__ cmaxg, int k,
    uint32_t* __restrict__ labels, double* __restrict__ psum, uint32_t* __restrict__ pcount,
    KmState* __restrict__ st) {
  double* acc = reinterpret_cast<double*>(smem);
  // stage starts on a 16-byte boundary (k * d rounded up to even)
  float* stage = reinterpret_cast<float*>(acc + (static_cast<size_t>(k) * d + 1) / 2 * 2);
  uint32_t* lab = reinterpret_cast<uint32_t*>(stage + static_cast<size_t>(kChunkRows) * ld);
  uint32_t* cnt = lab + kChunkRows;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int col = lane & 31;   // point inside the wave's 32 rows
  const int half = lane >> 5;
  const int ks_n = ld / 16;
  const int kb = (k + 31) / 32;
  const int64_t r0 = b * kBlockRows;
  const int rows_here = static_cast<int>(min<int64_t>(kBlockRows, n - r0));

  for (int e = tid; e < k * d; e += NT) acc[e] = 0.0;
  for (int e = tid; e < k; e += NT) cnt[e] = 0u;

  // A fragments: centroid 32 cb + col, dims 16 ks + 8 half .. +7
  bf16x8 ca[kMaxCB][kMaxKS];
  float c2v[kMaxCB][16];
#pragma unroll
  for (int cb = 0; cb < kMaxCB; ++cb) {
#pragma unroll
    for (int ks = 0; ks < kMaxKS; ++ks)
      ca[cb][ks] = (cb < kb && ks < ks_n)
                       ? load_frag(cb16 + static_cast<int64_t>(32 * cb + col) * ld + 16 * ks + 8 * half)
                       : bf16x8{};
    // output rows of this lane: centroid 32 cb + 8 (v / 4) + 4 half + v % 4
#pragma unroll
    for (int v = 0; v < 16; ++v)
      c2v[cb][v] = cb < kb ? c2g[32 * cb + 8 * (v >> 2) + 4 * half + (v & 3)] : INFINITY;
  }
  const float cmax = cmaxg[0];
  const float c2max = cmaxg[1];

  for (int c0 = 0; c0 < rows_here; c0 += kChunkRows) {
    const int nr = min(kChunkRows, rows_here - c0);
    __syncthreads();
    {
      const float* src = rows + (r0 + c0) * ld;
      const int nvec = nr * ld / 4;
      const f32x4* s4 = reinterpret_cast<const f32x4*>(src);
      f32x4* d4 = reinterpret_cast<f32x4*>(stage);
      for (int e = tid; e < nvec; e += NT) d4[e] = s4[e];
    }
    __syncthreads();

    // ---- certified assignment of this wave's 32 rows ----
    const int pr = 32 * wave + col;  // row inside the chunk
    const bool pvalid = pr < nr;
    const float* xr = stage + static_cast<size_t>(pvalid ? pr : 0) * ld;
    bf16x8 xb[kMaxKS];
    float x2p = 0.0f;
    bool finite = true;
#pragma unroll
    for (int ks = 0; ks < kMaxKS; ++ks) {
      if (ks < ks_n) {
        const f32x4 lo = *reinterpret_cast<const f32x4*>(xr + 16 * ks + 8 * half);
        const f32x4 hi = *reinterpret_cast<const f32x4*>(xr + 16 * ks + 8 * half + 4);
        u32x4 pk;
        pk[0] = bf16_bits_rne(lo[0]) | (bf16_bits_rne(lo[1]) << 16);
        pk[1] = bf16_bits_rne(lo[2]) | (bf16_bits_rne(lo[3]) << 16);
        pk[2] = bf16_bits_rne(hi[0]) | (bf16_bits_rne(hi[1]) << 16);
        pk[3] = bf16_bits_rne(hi[2]) | (bf16_bits_rne(hi[3]) << 16);
        xb[ks] = __builtin_bit_cast(bf16x8, pk);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          x2p = fmaf(lo[t], lo[t], x2p);
          x2p = fmaf(hi[t], hi[t], x2p);
          finite = finite && isfinite(lo[t]) && isfinite(hi[t]);
        }
      } else {
        xb[ks] = bf16x8{};
      }
    }
    const float x2 = x2p + __shfl_xor(x2p, 32);
    finite = finite && (__shfl_xor(finite ? 1 : 0, 32) != 0);

    f32x16 cacc[kMaxCB];
#pragma unroll
    for (int cb = 0; cb < kMaxCB; ++cb) cacc[cb] = f32x16{};
#pragma unroll
    for (int ks = 0; ks < kMaxKS; ++ks) {
      if (ks < ks_n) {
#pragma unroll
        for (int cb = 0; cb < kMaxCB; ++cb)
          if (cb < kb) cacc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ca[cb][ks], xb[ks], cacc[cb], 0, 0, 0);
      }
    }
    float amin = INFINITY;
    uint32_t aidx = kUncert;
#pragma unroll
    for (int cb = 0; cb < kMaxCB; ++cb) {
      if (cb < kb) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const float a = fmaf(-2.0f, cacc[cb][v], c2v[cb][v]);
          cacc[cb][v] = a;
          lexmin_f(amin, aidx, a, static_cast<uint32_t>(32 * cb + 8 * (v >> 2) + 4 * half + (v & 3)));
        }
      }
    }
    {
      const float v2 = __shfl_xor(amin, 32);
      const uint32_t i2 = __shfl_xor(aidx, 32);
      lexmin_f(amin, aidx, v2, i2);
    }
    // E >= |(s_c - ||x||^2) - a_c| for every c
    const double x2u = static_cast<double>(x2) * (1.0 + 3.814697265625e-06) + 1e-300;
    const double E = 2.0 * kC1 * sqrt(x2u) * static_cast<double>(cmax) * (1.0 + 2.384185791015625e-07) +
                     kKappa * (x2u + static_cast<double>(c2max));
    const float thr = __double2float_ru(static_cast<double>(amin) + 2.0 * E * (1.0 + 9.765625e-04));
    int ncand = 0;
#pragma unroll
    for (int cb = 0; cb < kMaxCB; ++cb) {
      if (cb < kb) {
#pragma unroll
        for (int v = 0; v < 16; ++v) ncand += cacc[cb][v] <= thr ? 1 : 0;
      }
    }
    ncand += __shfl_xor(ncand, 32);
    const bool certified = finite && ncand == 1 && amin < INFINITY && x2 < 1e30f;
    if (half == 0 && pvalid) lab[pr] = certified ? aidx : kUncert;
    // ---- FP64 assignment of uncertified rows: lanes = centroids ----
    uint64_t um = __ballot(half == 0 && pvalid && !certified);
    while (um) {
      const int p = __builtin_ctzll(um);
      um &= um - 1;
      const float* x = stage + static_cast<size_t>(32 * wave + p) * ld;
      double bv = INFINITY;
      uint32_t bi = kUncert;
      for (int c = lane; c < k; c += 64) {
        const double* cc = cent + static_cast<int64_t>(c) * ld;
        double s = 0.0;
        for (int j = 0; j < ld; ++j) s = s + Term<double, kEuclid>::f(static_cast<double>(x[j]), cc[j], 2.0);
        lexmin_d(bv, bi, s, static_cast<uint32_t>(c));
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double v2 = __shfl_xor(bv, off);
        const uint32_t i2 = __shfl_xor(bi, off);
        lexmin_d(bv, bi, v2, i2);
      }
      if (lane == 0) {
        // kmeans.rs:334-347: min starts at f64::MAX; nothing below it keeps 0
        lab[32 * wave + p] = bv < DBL_MAX ? bi : 0u;
        atomicAdd(&st->fallback, 1ull);
      }
    }
    if (half == 0 && pvalid) labels[r0 + c0 + pr] = lab[pr];
    __syncthreads();

    // ---- per-cluster sums, owner wave per cluster, rows in order ----
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
  __syncthreads();
}

__global__ __launch_bounds__(NT) void km_mixed_step_kernel(
    const float* __restrict__ rows, int64_t n, int ld, int d, const double* __restrict__ cent,
    const uint16_t* __restrict__ cb16, const float* __restrict__ c2g,
    const float* __restrict__ cmaxg, int k, uint32_t* __restrict__ labels,
    double* __restrict__ psum, uint32_t* __restrict__ pcount, KmState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  km_block_step(blockIdx.x, smem, rows, n, ld, d, cent, cb16, c2g, cmaxg, k, labels, psum, pcount,
                st);
}

size_t mixed_lds_bytes(const sc_index* ix, int k) {
  return sizeof(double) * ((static_cast<size_t>(k) * ix->d + 1) / 2 * 2) +
         sizeof(float) * static_cast<size_t>(kChunkRows) * ix->ld +
         sizeof(uint32_t) * (kChunkRows + static_cast<size_t>(k)) + 16;
}

}  // namespace

bool kmeans_mixed_supported(const sc_index* ix, int k) {
  return ix->dtype == SC_F32 && k >= 1 && k <= 32 * kMaxCB && ix->ld <= 16 * kMaxKS &&
         static_cast<int64_t>(k) * static_cast<int64_t>(ix->d) <= kMaxLdsAcc &&
         mixed_lds_bytes(ix, k) <= 96 * 1024;
}

int kmeans_mixed_step_launch(const sc_index* ix, KmWork& w, int k, const double* d_cent,
                             KmState* st, hipStream_t s) {
  const int ld = static_cast<int>(ix->ld);
  const int kpad = (k + 31) / 32 * 32;
  // per-iteration centroid terms (stream-ordered scratch)
  DevBuf terms;
  SC_HIP(terms.alloc(sizeof(uint16_t) * kpad * ld + sizeof(float) * (kpad + 4), s));
  uint16_t* cb16 = terms.as<uint16_t>();
  float* c2 = reinterpret_cast<float*>(cb16 + static_cast<size_t>(kpad) * ld);
  float* cmax = c2 + kpad;
  hipLaunchKernelGGL(km_prep_centroids_kernel, dim3(1), dim3(NT), 0, s, d_cent, k, kpad, ld, cb16,
                     c2, cmax, st);
  SC_HIP(hipGetLastError());
  if (w.nblocks == 0) return SC_OK;
  ProfScope ps("kmeans_mixed_assign", s);
  hipLaunchKernelGGL(km_mixed_step_kernel, dim3(static_cast<unsigned>(w.nblocks)), dim3(NT),
                     mixed_lds_bytes(ix, k), s, static_cast<const float*>(ix->rows),
                     static_cast<int64_t>(ix->n), ld, static_cast<int>(ix->d), d_cent, cb16, c2,
                     cmax, k, w.labels, w.psum, w.pcount, st);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// ------------------------------------------------------------------ batch
// Many independent problems (uneven n, d, k) in one persistent kernel per
// Lloyd iteration.  Work items are (problem, block) pairs; a fixed grid of
// workgroups strides over them.  The workgroup that finishes the last block
// of a problem (a per-problem arrival counter) reduces that problem's block
// partials in the same fixed order as the single-problem path, applies the
// distortion, centroid update and stop rule, and prepares the next
// iteration's BF16 centroids.  Every problem therefore gets bit-identical
// results to sc_kmeans_run(SC_MIXED) on it alone.  The iterations are
// captured in one hipGraph; the grid size is picked by timing the first
// iterations (any choice gives the same results).
namespace {

struct KmbProblem {
  int64_t n;
  int d, ld, k, kpad;
  int64_t nb, ng;
  int64_t row_off, cent_off, cb_off, c2_off, lab_off, ps_off, pc_off, gs_off, sum_off, cnt_off,
      x2_off;
};

struct KmbArena {
  float* rows;
  double* cent;
  uint16_t* cb;
  float* c2;  // per problem: c2[kpad] | max ||c|| | max ||c||^2
  uint32_t* lab;
  double* ps;
  uint32_t* pc;
  double* gs;
  double* sums;
  uint64_t* cnt;
  double* x2;  // per problem: x2 per block
  KmState* st;
  uint32_t* arrivals;
};

__global__ void kmb_init_kernel(KmbArena a, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  KmState* st = a.st + i;
  st->distortion = DBL_MAX;
  st->dist = 0.0;
  st->x2total = 0.0;
  st->iters = 0;
  st->done = 0;
  st->fallback = 0;
  a.arrivals[i] = 0u;
}

// ||x||^2 per (problem, block): same order as the single-problem x2 kernel.
__global__ __launch_bounds__(NT) void kmb_x2_kernel(const KmbProblem* __restrict__ P,
                                                    const uint2* __restrict__ items, int64_t nitems,
                                                    KmbArena a) {
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const KmbProblem p = P[items[it].x];
    const int64_t b = items[it].y;
    const float* rows = a.rows + p.row_off;
    const int64_t r0 = b * kBlockRows;
    double s = 0.0;
    for (int j = 0; j < kBlockRows / NT; ++j) {
      const int64_t r = r0 + tid * (kBlockRows / NT) + j;
      if (r < p.n) {
        double t = 0.0;
        for (int c = 0; c < p.ld; ++c) {
          const double v = static_cast<double>(rows[r * p.ld + c]);
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
    if (tid == 0) a.x2[p.x2_off + b] = red[0];
    __syncthreads();
  }
}

__global__ void kmb_x2_total_kernel(const KmbProblem* __restrict__ P, int count, KmbArena a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const KmbProblem p = P[i];
  double s = 0.0;
  for (int64_t b = 0; b < p.nb; ++b) s = s + a.x2[p.x2_off + b];
  a.st[i].x2total = s;
}

__global__ __launch_bounds__(NT) void kmb_prep_kernel(const KmbProblem* __restrict__ P, KmbArena a) {
  __shared__ double red[NT];
  const int i = blockIdx.x;
  if (a.st[i].done) return;
  const KmbProblem p = P[i];
  km_prep_centroids_body(a.cent + p.cent_off, p.k, p.kpad, p.ld, a.cb + p.cb_off, a.c2 + p.c2_off,
                         a.c2 + p.c2_off + p.kpad, red);
}

// Reduction, distortion, update, stop rule and next-iteration centroid terms
// of problem i, by one workgroup, in the single-problem path's order.
__device__ void kmb_finish(const KmbProblem& p, int i, const KmbArena& a, double* red,
                           double* term) {
  const int tid = threadIdx.x;
  const int k = p.k, d = p.d;
  const int64_t kd = static_cast<int64_t>(k) * d;
  const volatile double* ps = a.ps + p.ps_off;
  const volatile uint32_t* pc = a.pc + p.pc_off;
  double* gs = a.gs + p.gs_off;
  for (int64_t t = tid; t < p.ng * kd; t += NT) {
    const int64_t g = t / kd, e = t % kd;
    const int64_t b0 = g * kGroupBlocks, b1 = min<int64_t>(p.nb, b0 + kGroupBlocks);
    double s = 0.0;
    for (int64_t bb = b0; bb < b1; ++bb) s = s + ps[bb * kd + e];
    gs[g * kd + e] = s;
  }
  __syncthreads();
  const volatile double* gsv = gs;
  double* sums = a.sums + p.sum_off;
  uint64_t* cnt = a.cnt + p.cnt_off;
  for (int64_t e = tid; e < kd; e += NT) {
    double s = 0.0;
    for (int64_t g = 0; g < p.ng; ++g) s = s + gsv[g * kd + e];
    sums[e] = s;
  }
  for (int c = tid; c < k; c += NT) {
    uint64_t n = 0;
    for (int64_t bb = 0; bb < p.nb; ++bb) n += pc[bb * k + c];
    cnt[c] = n;
  }
  __syncthreads();
  const volatile double* sv = sums;
  const volatile uint64_t* cv = cnt;
  double* cent = a.cent + p.cent_off;
  for (int c = tid; c < k; c += NT) {
    const double* cc = cent + static_cast<int64_t>(c) * p.ld;
    double dot = 0.0, c2 = 0.0;
    for (int j = 0; j < d; ++j) {
      dot = dot + cc[j] * sv[static_cast<int64_t>(c) * d + j];
      c2 = c2 + cc[j] * cc[j];
    }
    term[c] = cv[c] > 0 ? static_cast<double>(cv[c]) * c2 - 2.0 * dot : 0.0;
  }
  __syncthreads();
  KmState* st = a.st + i;
  if (tid == 0) {
    double s = st->x2total;
    for (int c = 0; c < k; ++c) s = s + term[c];
    st->dist = s;
  }
  __syncthreads();
  for (int64_t e = tid; e < kd; e += NT) {
    const int c = static_cast<int>(e / d), j = static_cast<int>(e % d);
    if (cv[c] > 0) cent[static_cast<int64_t>(c) * p.ld + j] = sv[e] / static_cast<double>(cv[c]);
  }
  __shared__ int s_done;
  if (tid == 0) {
    st->iters += 1;
    if (st->distortion <= st->dist) {
      st->done = 1;
    } else {
      st->distortion = st->dist;
    }
    s_done = static_cast<int>(st->done);
    a.arrivals[i] = 0u;
  }
  __syncthreads();
  if (!s_done)
    km_prep_centroids_body(cent, k, p.kpad, p.ld, a.cb + p.cb_off, a.c2 + p.c2_off,
                           a.c2 + p.c2_off + p.kpad, red);
}

__global__ __launch_bounds__(NT) void kmb_step_kernel(const KmbProblem* __restrict__ P,
                                                      const uint2* __restrict__ items,
                                                      int64_t nitems, KmbArena a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NT];
  __shared__ double term[32 * kMaxCB];
  __shared__ int s_last;
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int i = static_cast<int>(items[it].x);
    const int64_t b = items[it].y;
    KmState* st = a.st + i;
    // every block of problem i in this launch sees the same value: the
    // flag only changes in the finisher, after all of them have arrived
    if (*reinterpret_cast<volatile uint32_t*>(&st->done)) continue;
    const KmbProblem p = P[i];
    km_block_step(b, smem, a.rows + p.row_off, p.n, p.ld, p.d, a.cent + p.cent_off,
                  a.cb + p.cb_off, a.c2 + p.c2_off, a.c2 + p.c2_off + p.kpad, p.k,
                  a.lab + p.lab_off, a.ps + p.ps_off, a.pc + p.pc_off, st);
    if (threadIdx.x == 0) {
      __threadfence();
      const uint32_t old = atomicAdd(&a.arrivals[i], 1u);
      s_last = old == static_cast<uint32_t>(p.nb - 1) ? 1 : 0;
    }
    __syncthreads();
    if (s_last) {
      __threadfence();
      kmb_finish(p, i, a, red, term);
    }
    __syncthreads();
  }
}

size_t kmb_lds_bytes(int k, int d, int ld) {
  return sizeof(double) * ((static_cast<size_t>(k) * d + 1) / 2 * 2) +
         sizeof(float) * static_cast<size_t>(kChunkRows) * ld +
         sizeof(uint32_t) * (kChunkRows + static_cast<size_t>(k)) + 16;
}

struct TuneEntry {
  int64_t key;
  unsigned grid;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;

}  // namespace

}  // namespace sc

using namespace sc;

extern "C" int sc_kmeans_batch_run(const sc_kmeans_problem* problems, size_t count, sc_dtype dtype,
                                   size_t max_iter, sc_mode mode, double* centroids_out,
                                   uint32_t* labels_out, uint64_t* sizes_out,
                                   double* distortion_out, uint32_t* iters_out) {
  clear_error();
  if (count == 0) return SC_OK;
  if (!problems || !centroids_out || !labels_out || !sizes_out || !distortion_out || !iters_out)
    return fail(SC_PARAMETERS_ERROR, "null argument");
  if (max_iter == 0)
    return fail(SC_FIT_FAILED, "invalid maximum number of iterations: " + std::to_string(max_iter));
  if (dtype != SC_F64 && dtype != SC_F32) return fail(SC_ERR_UNSUPPORTED, "dtype must be f64 or f32");
  for (size_t i = 0; i < count; ++i) {
    if (problems[i].k < 2)
      return fail(SC_FIT_FAILED, "invalid number of clusters: " + std::to_string(problems[i].k));
    if (problems[i].n == 0 || !problems[i].rows || !problems[i].init_centroids)
      return fail(SC_PARAMETERS_ERROR, "problem " + std::to_string(i) + " has no rows");
  }
  int dev = 0;
  SC_HIP(hipGetDevice(&dev));
  hipStream_t s = stream_for_device(dev);

  bool batched = mode == SC_MIXED && dtype == SC_F32;
  size_t lds = 0;
  for (size_t i = 0; i < count && batched; ++i) {
    const int ld = static_cast<int>(padded_dim(problems[i].d));
    const int k = static_cast<int>(problems[i].k);
    lds = std::max(lds, kmb_lds_bytes(k, static_cast<int>(problems[i].d), ld));
    if (k > 32 * kMaxCB || ld > 16 * kMaxKS ||
        static_cast<int64_t>(k) * static_cast<int64_t>(problems[i].d) > kMaxLdsAcc)
      batched = false;
  }
  if (lds > 96 * 1024) batched = false;

  if (!batched) {
    // the FP64 form, one problem at a time (the comparator of the batch)
    size_t co = 0, lo = 0, so = 0;
    for (size_t i = 0; i < count; ++i) {
      const sc_kmeans_problem& pr = problems[i];
      sc_index* ix = nullptr;
      if (int rc = sc_index_create(pr.rows, dtype, pr.n, pr.d, &ix)) return rc;
      std::copy(pr.init_centroids, pr.init_centroids + pr.k * pr.d, centroids_out + co);
      const int rc = sc_kmeans_run(ix, pr.k, max_iter, mode, centroids_out + co, labels_out + lo,
                                   sizes_out + so, distortion_out + i, iters_out + i);
      sc_index_destroy(ix);
      if (rc) return rc;
      co += pr.k * pr.d;
      lo += pr.n;
      so += pr.k;
    }
    return SC_OK;
  }

  // ---- layout ----
  std::vector<KmbProblem> P(count);
  int64_t rows_e = 0, cent_e = 0, cb_e = 0, c2_e = 0, lab_e = 0, ps_e = 0, pc_e = 0, gs_e = 0,
          sum_e = 0, cnt_e = 0, x2_e = 0;
  std::vector<uint2> items;
  for (size_t i = 0; i < count; ++i) {
    const sc_kmeans_problem& pr = problems[i];
    KmbProblem& p = P[i];
    p.n = static_cast<int64_t>(pr.n);
    p.d = static_cast<int>(pr.d);
    p.ld = static_cast<int>(padded_dim(pr.d));
    p.k = static_cast<int>(pr.k);
    p.kpad = (p.k + 31) / 32 * 32;
    p.nb = (p.n + kBlockRows - 1) / kBlockRows;
    p.ng = (p.nb + kGroupBlocks - 1) / kGroupBlocks;
    const int64_t kd = static_cast<int64_t>(p.k) * p.d;
    p.row_off = rows_e;   rows_e += (p.n * p.ld + 3) / 4 * 4;
    p.cent_off = cent_e;  cent_e += static_cast<int64_t>(p.k) * p.ld;
    p.cb_off = cb_e;      cb_e += (static_cast<int64_t>(p.kpad) * p.ld + 7) / 8 * 8;
    p.c2_off = c2_e;      c2_e += p.kpad + 4;
    p.lab_off = lab_e;    lab_e += p.n;
    p.ps_off = ps_e;      ps_e += p.nb * kd;
    p.pc_off = pc_e;      pc_e += p.nb * p.k;
    p.gs_off = gs_e;      gs_e += p.ng * kd;
    p.sum_off = sum_e;    sum_e += kd;
    p.cnt_off = cnt_e;    cnt_e += p.k;
    p.x2_off = x2_e;      x2_e += p.nb;
  }
  // items, block-major across problems so concurrent workgroups spread over
  // problems and each problem's last block arrives early
  int64_t maxnb = 0;
  for (auto& p : P) maxnb = std::max(maxnb, p.nb);
  for (int64_t b = 0; b < maxnb; ++b)
    for (size_t i = 0; i < count; ++i)
      if (b < P[i].nb) items.push_back(make_uint2(static_cast<unsigned>(i), static_cast<unsigned>(b)));
  const int64_t nitems = static_cast<int64_t>(items.size());

  DevBuf bP, bI, bRows, bCent, bCb, bC2, bLab, bPs, bPc, bGs, bSum, bCnt, bX2, bSt, bArr;
  SC_HIP(bP.alloc(sizeof(KmbProblem) * count, s));
  SC_HIP(bI.alloc(sizeof(uint2) * nitems, s));
  SC_HIP(bRows.alloc(sizeof(float) * rows_e, s));
  SC_HIP(bCent.alloc(sizeof(double) * cent_e, s));
  SC_HIP(bCb.alloc(sizeof(uint16_t) * cb_e, s));
  SC_HIP(bC2.alloc(sizeof(float) * c2_e, s));
  SC_HIP(bLab.alloc(sizeof(uint32_t) * lab_e, s));
  SC_HIP(bPs.alloc(sizeof(double) * std::max<int64_t>(ps_e, 1), s));
  SC_HIP(bPc.alloc(sizeof(uint32_t) * std::max<int64_t>(pc_e, 1), s));
  SC_HIP(bGs.alloc(sizeof(double) * std::max<int64_t>(gs_e, 1), s));
  SC_HIP(bSum.alloc(sizeof(double) * std::max<int64_t>(sum_e, 1), s));
  SC_HIP(bCnt.alloc(sizeof(uint64_t) * std::max<int64_t>(cnt_e, 1), s));
  SC_HIP(bX2.alloc(sizeof(double) * std::max<int64_t>(x2_e, 1), s));
  SC_HIP(bSt.alloc(sizeof(KmState) * count, s));
  SC_HIP(bArr.alloc(sizeof(uint32_t) * count, s));
  KmbArena a{bRows.as<float>(), bCent.as<double>(), bCb.as<uint16_t>(), bC2.as<float>(),
             bLab.as<uint32_t>(), bPs.as<double>(), bPc.as<uint32_t>(), bGs.as<double>(),
             bSum.as<double>(), bCnt.as<uint64_t>(), bX2.as<double>(), bSt.as<KmState>(),
             bArr.as<uint32_t>()};
  SC_HIP(hipMemcpyAsync(bP.p, P.data(), sizeof(KmbProblem) * count, hipMemcpyHostToDevice, s));
  SC_HIP(hipMemcpyAsync(bI.p, items.data(), sizeof(uint2) * nitems, hipMemcpyHostToDevice, s));
  SC_HIP(hipMemsetAsync(bCent.p, 0, sizeof(double) * cent_e, s));
  for (size_t i = 0; i < count; ++i) {
    const sc_kmeans_problem& pr = problems[i];
    if (int rc = upload_padded(pr.rows, SC_F32, pr.n, pr.d, P[i].ld, a.rows + P[i].row_off, s))
      return rc;
    SC_HIP(hipMemcpy2DAsync(a.cent + P[i].cent_off, P[i].ld * sizeof(double), pr.init_centroids,
                            pr.d * sizeof(double), pr.d * sizeof(double), pr.k,
                            hipMemcpyHostToDevice, s));
  }
  const KmbProblem* dP = bP.as<KmbProblem>();
  const uint2* dI = bI.as<uint2>();
  const unsigned cgrid = static_cast<unsigned>((count + 255) / 256);
  hipLaunchKernelGGL(kmb_init_kernel, dim3(cgrid), dim3(256), 0, s, a, static_cast<int>(count));
  hipLaunchKernelGGL(kmb_x2_kernel, dim3(1024), dim3(NT), 0, s, dP, dI, nitems, a);
  hipLaunchKernelGGL(kmb_x2_total_kernel, dim3(cgrid), dim3(256), 0, s, dP, static_cast<int>(count), a);
  hipLaunchKernelGGL(kmb_prep_kernel, dim3(static_cast<unsigned>(count)), dim3(NT), 0, s, dP, a);
  SC_HIP(hipGetLastError());

  // ---- grid size: cached, or timed on the first iterations ----
  const int64_t key = static_cast<int64_t>(count) * 1000003 + nitems;
  unsigned grid = 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& e : g_tune)
      if (e.key == key) grid = e.grid;
  }
  size_t it_done = 0;
  if (grid == 0) {
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
    const unsigned cand[3] = {static_cast<unsigned>(cus), static_cast<unsigned>(2 * cus),
                              static_cast<unsigned>(4 * cus)};
    float best = 1e30f;
    hipEvent_t e0, e1;
    SC_HIP(hipEventCreate(&e0));
    SC_HIP(hipEventCreate(&e1));
    for (int c = 0; c < 3 && it_done < max_iter; ++c) {
      SC_HIP(hipEventRecord(e0, s));
      hipLaunchKernelGGL(kmb_step_kernel, dim3(cand[c]), dim3(NT), lds, s, dP, dI, nitems, a);
      SC_HIP(hipGetLastError());
      SC_HIP(hipEventRecord(e1, s));
      SC_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      SC_HIP(hipEventElapsedTime(&ms, e0, e1));
      ++it_done;
      if (ms < best) {
        best = ms;
        grid = cand[c];
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.push_back(TuneEntry{key, grid});
  }
  // ---- remaining iterations as one graph ----
  if (it_done < max_iter) {
    ProfScope ps("kmeans_batch", s);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (size_t it = it_done; it < max_iter; ++it)
      hipLaunchKernelGGL(kmb_step_kernel, dim3(grid), dim3(NT), lds, s, dP, dI, nitems, a);
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ce != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return hip_fail(ce, "hipGraphInstantiate");
    }
    ce = hipGraphLaunch(exec, s);
    const hipError_t se = hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (ce != hipSuccess) return hip_fail(ce, "hipGraphLaunch");
    if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize");
  }
  // ---- results ----
  std::vector<KmState> hs(count);
  SC_HIP(hipMemcpyAsync(hs.data(), bSt.p, sizeof(KmState) * count, hipMemcpyDeviceToHost, s));
  size_t co = 0, lo = 0, so = 0;
  for (size_t i = 0; i < count; ++i) {
    const KmbProblem& p = P[i];
    SC_HIP(hipMemcpy2DAsync(centroids_out + co, p.d * sizeof(double), a.cent + p.cent_off,
                            p.ld * sizeof(double), p.d * sizeof(double), p.k,
                            hipMemcpyDeviceToHost, s));
    SC_HIP(hipMemcpyAsync(labels_out + lo, a.lab + p.lab_off, sizeof(uint32_t) * p.n,
                          hipMemcpyDeviceToHost, s));
    SC_HIP(hipMemcpyAsync(sizes_out + so, a.cnt + p.cnt_off, sizeof(uint64_t) * p.k,
                          hipMemcpyDeviceToHost, s));
    co += static_cast<size_t>(p.k) * p.d;
    lo += p.n;
    so += p.k;
  }
  SC_HIP(hipStreamSynchronize(s));
  uint64_t fb = 0;
  for (size_t i = 0; i < count; ++i) {
    distortion_out[i] = hs[i].distortion;
    iters_out[i] = hs[i].iters;
    fb += hs[i].fallback;
  }
  sc_stats& st = last_stats();
  st = sc_stats{};
  st.queries = count;
  st.uncertified = fb;
  return SC_OK;
}
