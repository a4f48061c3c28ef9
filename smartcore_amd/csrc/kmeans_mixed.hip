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
__global__ __launch_bounds__(NT) void km_prep_centroids_kernel(const double* __restrict__ cent,
                                                               int k, int kpad, int ld,
                                                               uint16_t* __restrict__ cb16,
                                                               float* __restrict__ c2,
                                                               float* __restrict__ cmax,
                                                               const KmState* __restrict__ st) {
  if (st->done) return;
  __shared__ double red[NT];
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
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// Dynamic LDS: acc[k*d] f64 | stage[kChunkRows*ld] f32 | lab[kChunkRows] | cnt[k].
__global__ __launch_bounds__(NT) void km_mixed_step_kernel(
    const float* __restrict__ rows, int64_t n, int ld, int d, const double* __restrict__ cent,
    const uint16_t* __restrict__ cb16, const float* __restrict__ c2g,
    const float* __restrict__ cmaxg, int k, uint32_t* __restrict__ labels,
    double* __restrict__ psum, uint32_t* __restrict__ pcount, KmState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
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
  const int64_t b = blockIdx.x;
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

}  // namespace sc
