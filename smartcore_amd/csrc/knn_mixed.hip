// Mixed-precision Euclidean kNN (SC_MIXED) for f32 rows: BF16 MFMA cross
// terms select candidates under a rigorous error bound; an FP64 replay in
// the reference's arithmetic reproduces LinearKNNSearch::find
// (linear_search.rs:52-84) exactly -- same indices, same heap-array order,
// same distance bits -- as the SC_PLAIN_FP64 form.
//
// Bound.  Let s be the reference's squared distance (f32 differences and
// squares, f64 sum; euclidian.rs:57-62), x2 = ||x||^2, y2 = ||y||^2 and
// C = bf16(x) . bf16(y) accumulated in FP32 by the MFMA.  With u = 2^-8
// (bf16 unit roundoff) the cross-term error is at most
//   |x.y - C| <= (2u + u^2 + 128 * 2^-22) ||x|| ||y|| = c1 ||x|| ||y||,
// the last term allowing four FP32 ulps per accumulation step.  The f32
// rounding inside s and every FP32 rounding of the epilogue are covered by
// a relative margin of 2^-18 (x2 + y2).  Hence
//   L = (1 - 2^-18)(x2 + y2) - 2C - 2 c1 ||x|| ||y||  <=  s  <=
//   U = (1 + 2^-16)(x2 + y2) - 2C + 2 c1 ||x|| ||y||.
//
// Pass 1 (scan, MFMA).  A workgroup of four waves owns 512 queries (128 per
// wave, held as BF16 B-fragments in registers) and one contiguous segment
// of rows, streamed 32 rows at a time in index order through
// v_mfma_f32_32x32x16_bf16.  Per query the wave keeps the k smallest U seen
// in the segment; every row whose L is below that running k-th U is
// appended (index, L) to the query's list for the segment, in index order.
// A row that enters the reference heap has s below the running k-th
// smallest s of all earlier rows, which is at most the running k-th U of
// the segment, so the lists are a superset of the heap entrants.
// Pass 2 (replay, FP64).  One thread per query walks its lists in index
// order and rescores a candidate exactly only if its L is below the s of
// the current heap root (otherwise it cannot enter), then applies the heap
// rule.  Rows that would not enter the reference heap are no-ops, so the
// final heap array equals the reference's.  Queries whose lists overflow,
// or whose rows are not finite / too large for the bound, are recomputed
// by the sequential exact kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "metric.hpp"

namespace sc {

// from knn_exact.hip
int launch_serial_exact_f32(const sc_index* index, const float* d_queries, int k,
                            const uint32_t* list, const uint32_t* list_n, int64_t* d_idx,
                            double* d_dist, uint32_t* d_count, hipStream_t stream);

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kQPW = 128;                 // queries per wave (4 column blocks)
constexpr int kQPB = kWaves * kQPW;       // queries per workgroup
constexpr int kRows = 32;                 // rows per tile
constexpr int kMaxKS = 8;                 // K steps of 16 -> dimension <= 128
constexpr int kMaxK = 32;                 // largest k served by this form
constexpr double kC1 = 0.0078125 + 1.52587890625e-05 + 128.0 * 2.384185791015625e-07;
constexpr double kLoMargin = 3.814697265625e-06;   // 2^-18
constexpr double kHiMargin = 1.52587890625e-05;    // 2^-16
constexpr double kNormLimit = 1e30;                // squared norms beyond this use FP64

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Per row: BF16 copy, y2m = RD((1 - 2^-18) y2), nyu = RU(||y|| (1 + 2^-20)),
// y2n = RN(y2).  A row whose values are not finite or whose squared norm
// reaches kNormLimit marks `bad` (the whole index then uses FP64).
__global__ void prep_rows_kernel(const float* __restrict__ rows, int64_t n, int ld,
                                 uint16_t* __restrict__ out_b, float* __restrict__ y2m,
                                 float* __restrict__ nyu, uint32_t* __restrict__ bad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rows + i * ld;
  uint16_t* o = out_b + i * ld;
  double s = 0.0;
  bool ok = true;
  for (int j = 0; j < ld; ++j) {
    const float v = r[j];
    ok = ok && isfinite(v);
    s += static_cast<double>(v) * static_cast<double>(v);
    o[j] = f32_to_bf16_rne(v);
  }
  if (!ok || !(s < kNormLimit)) {
    atomicOr(bad, 1u);
    y2m[i] = INFINITY;
    nyu[i] = INFINITY;
    return;
  }
  y2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  nyu[i] = __double2float_ru(sqrt(s) * (1.0 + 9.5367431640625e-07));
}

__global__ void fill_f32_kernel(float* p, int64_t n, float v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per query: BF16 copy, x2m, cx = RU(2 c1 ||x|| (1 + 2^-20)); a query that
// is not finite or too large is listed for the exact kernel.
__global__ void prep_queries_kernel(const float* __restrict__ qs, int64_t q, int ld,
                                    uint16_t* __restrict__ out_b, float* __restrict__ x2m,
                                    float* __restrict__ cx, uint32_t* __restrict__ list,
                                    uint32_t* __restrict__ list_n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= q) return;
  const float* r = qs + i * ld;
  uint16_t* o = out_b + i * ld;
  double s = 0.0;
  bool ok = true;
  for (int j = 0; j < ld; ++j) {
    const float v = r[j];
    ok = ok && isfinite(v);
    s += static_cast<double>(v) * static_cast<double>(v);
    o[j] = f32_to_bf16_rne(v);
  }
  if (!ok || !(s < kNormLimit)) {
    // never a candidate in pass 1; the exact kernel takes it
    x2m[i] = INFINITY;
    cx[i] = 0.0f;
    const uint32_t slot = atomicAdd(list_n, 1u);
    list[slot] = static_cast<uint32_t>(i);
    return;
  }
  x2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  cx[i] = __double2float_ru(2.0 * kC1 * sqrt(s) * (1.0 + 9.5367431640625e-07));
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

__global__ __launch_bounds__(256, 1) void mixed_scan_kernel(
    const uint16_t* __restrict__ rows_b, const float* __restrict__ y2m,
    const float* __restrict__ nyu, int64_t n, int ld, const uint16_t* __restrict__ q_b,
    const float* __restrict__ x2m, const float* __restrict__ cxq, int64_t q, int k,
    int64_t seg_len, int nseg, uint32_t* __restrict__ cand_idx, float* __restrict__ cand_L,
    uint32_t* __restrict__ cand_count, int cap) {
  __shared__ float topu[kQPB][kMaxK];  // per query: k smallest U, ascending
  __shared__ uint32_t ccount[kQPB];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 31;   // query column inside a 32-wide block
  const int half = lane >> 5;  // which 8-element K half / which row quad
  const int64_t qw0 = static_cast<int64_t>(blockIdx.x) * kQPB + wave * kQPW;
  const int seg = blockIdx.y;
  const int64_t r_begin = static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);
  const int ks_n = ld / 16;

  for (int e = threadIdx.x; e < kQPB * kMaxK; e += 256) (&topu[0][0])[e] = INFINITY;
  for (int e = threadIdx.x; e < kQPB; e += 256) ccount[e] = 0u;
  __syncthreads();

  // B fragments: block cb, K step ks: query (qw0 + 32 cb + col), dims
  // 16 ks + 8 half .. +7.  Missing queries read row 0 and never hit.
  bf16x8 bq[4][kMaxKS];
  float thr[4], cxv[4], x2v[4];
  bool qvalid[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int64_t qi = qw0 + 32 * cb + col;
    // queries past the batch, and queries prep listed for the exact kernel
    // (x2m = inf), never produce candidates
    qvalid[cb] = qi < q && x2m[qi < q ? qi : 0] < INFINITY;
    const int64_t qr = qi < q ? qi : 0;
#pragma unroll
    for (int ks = 0; ks < kMaxKS; ++ks)
      bq[cb][ks] = ks < ks_n ? load_frag(q_b + qr * ld + 16 * ks + 8 * half) : bf16x8{};
    cxv[cb] = qvalid[cb] ? cxq[qr] : 0.0f;
    x2v[cb] = qvalid[cb] ? x2m[qr] : INFINITY;
    thr[cb] = INFINITY;
  }

  for (int64_t t0 = r_begin; t0 < r_end; t0 += kRows) {
    // A fragments: row t0 + col, dims 16 ks + 8 half .. +7
    const int64_t arow = min<int64_t>(t0 + col, n - 1);
    bf16x8 a[kMaxKS];
#pragma unroll
    for (int ks = 0; ks < kMaxKS; ++ks)
      a[ks] = ks < ks_n ? load_frag(rows_b + arow * ld + 16 * ks + 8 * half) : bf16x8{};
    // per-row terms for this lane's 16 output rows: rows 8g + 4half + 0..3
    float yv[16], nv[16];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int64_t rb = t0 + 8 * g + 4 * half;
      const f32x4 y4 = *reinterpret_cast<const f32x4*>(y2m + rb);
      const f32x4 n4 = *reinterpret_cast<const f32x4*>(nyu + rb);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const bool live = rb + v < r_end;
        yv[4 * g + v] = live ? y4[v] : INFINITY;
        nv[4 * g + v] = n4[v];
      }
    }
    f32x16 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = f32x16{};
#pragma unroll
    for (int ks = 0; ks < kMaxKS; ++ks) {
      if (ks < ks_n) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], bq[cb][ks], acc[cb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      uint32_t hits = 0;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        float val = fmaf(-2.0f, acc[cb][v], yv[v]);
        val = fmaf(-cxv[cb], nv[v], val);
        hits |= (val < thr[cb] ? 1u : 0u) << v;
      }
      if (!qvalid[cb]) hits = 0;
      if (__ballot(hits != 0) == 0) continue;
      // Rare path: in row order, each lane holding row r for its query
      // inserts U into the query's running top-k and appends (row, L).
      const int ql = wave * kQPW + 32 * cb + col;  // local query slot
      const int64_t qi = qw0 + 32 * cb + col;
      for (int r = 0; r < kRows; ++r) {
        const int h = (r >> 2) & 1;
        const int v = (r >> 3) * 4 + (r & 3);
        if (half == h && ((hits >> v) & 1u)) {
          const double C = acc[cb][v];
          const double base = static_cast<double>(x2v[cb]) + static_cast<double>(yv[v]);
          const double e = static_cast<double>(cxv[cb]) * static_cast<double>(nv[v]);
          const double L = base - 2.0 * C - e;
          const double U = base * ((1.0 + kHiMargin) / (1.0 - kLoMargin)) - 2.0 * C + e;
          float* tk = topu[ql];
          const float uf = __double2float_ru(U);
          if (uf < tk[k - 1]) {
            int pos = k - 1;
            while (pos > 0 && tk[pos - 1] > uf) {
              tk[pos] = tk[pos - 1];
              --pos;
            }
            tk[pos] = uf;
          }
          const uint32_t c = ccount[ql];
          if (c < static_cast<uint32_t>(cap)) {
            const size_t slot = (static_cast<size_t>(qi) * nseg + seg) * cap + c;
            cand_idx[slot] = static_cast<uint32_t>(t0 + r);
            cand_L[slot] = __double2float_rd(L);
          }
          ccount[ql] = c + 1;
        }
      }
      // T - x2m rounded up (the f64 difference of two f32 values is exact)
      thr[cb] = __double2float_ru(static_cast<double>(topu[ql][k - 1]) -
                                  static_cast<double>(x2v[cb]));
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kQPB; e += 256) {
    const int64_t qi = static_cast<int64_t>(blockIdx.x) * kQPB + e;
    if (qi < q) cand_count[static_cast<size_t>(qi) * nseg + seg] = ccount[e];
  }
}

__device__ inline void sift_down_s(double* hd, int64_t* hi, double* hs, int kk, int n) {
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
    double ts = hs[kk];
    hs[kk] = hs[j];
    hs[j] = ts;
    kk = j;
  }
}

// Pass 2: exact replay over the candidate lists (see the file comment).
__global__ void mixed_replay_kernel(const float* __restrict__ rows, const float* __restrict__ qs,
                                    int ld, int64_t q, int k, int nseg, int cap,
                                    const uint32_t* __restrict__ cand_idx,
                                    const float* __restrict__ cand_L,
                                    const uint32_t* __restrict__ cand_count,
                                    const float* __restrict__ x2m, double* __restrict__ hs_all,
                                    int64_t* __restrict__ out_idx, double* __restrict__ out_dist,
                                    uint32_t* __restrict__ out_count, uint32_t* __restrict__ list,
                                    uint32_t* __restrict__ list_n,
                                    unsigned long long* __restrict__ counters) {
  const int64_t qi = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (qi >= q) return;
  if (!(x2m[qi] < INFINITY)) return;  // listed for the exact kernel already
  double* hd = out_dist + qi * k;
  int64_t* hi = out_idx + qi * k;
  double* hs = hs_all + qi * k;
  for (int j = 0; j < k; ++j) {
    hd[j] = INFINITY;
    hi[j] = -1;
    hs[j] = INFINITY;
  }
  for (int s = 0; s < nseg; ++s) {
    if (cand_count[static_cast<size_t>(qi) * nseg + s] > static_cast<uint32_t>(cap)) {
      const uint32_t slot = atomicAdd(list_n, 1u);
      list[slot] = static_cast<uint32_t>(qi);
      return;
    }
  }
  const float* qrow = qs + qi * ld;
  unsigned long long seen = 0, rescored = 0;
  for (int s = 0; s < nseg; ++s) {
    const uint32_t cnt = cand_count[static_cast<size_t>(qi) * nseg + s];
    const size_t base = (static_cast<size_t>(qi) * nseg + s) * cap;
    seen += cnt;
    for (uint32_t c = 0; c < cnt; ++c) {
      const float L = cand_L[base + c];
      if (!(static_cast<double>(L) < hs[0])) continue;
      const int64_t r = cand_idx[base + c];
      const float* rrow = rows + r * ld;
      double sum = 0.0;
      for (int j = 0; j < ld; ++j) sum = sum + Term<float, kEuclid>::f(qrow[j], rrow[j], 2.0);
      ++rescored;
      const double d = __builtin_sqrt(sum);
      if (d < hd[0]) {
        hd[0] = d;
        hi[0] = r;
        hs[0] = sum;
        if (k > 1) sift_down_s(hd, hi, hs, 0, k - 1);
      }
    }
  }
  atomicAdd(&counters[0], seen);
  atomicAdd(&counters[1], rescored);
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

// Builds the BF16 copy and row terms of an f32 index once.
int ensure_bf16(sc_index* ix, hipStream_t s, bool* bad_out) {
  std::lock_guard<std::mutex> lk(*static_cast<std::mutex*>(ix->mu));
  if (ix->rows_bf16) {
    *bad_out = ix->ldb == 0;
    return SC_OK;
  }
  const int64_t n = static_cast<int64_t>(ix->n);
  const int64_t npad = (n + kRows - 1) / kRows * kRows + kRows;
  void* b = nullptr;
  float* terms = nullptr;
  SC_HIP(hipMalloc(&b, std::max<size_t>(1, n) * ix->ld * sizeof(uint16_t)));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&terms), sizeof(float) * 2 * npad));
  DevBuf flag;
  SC_HIP(flag.alloc(4, s));
  SC_HIP(hipMemsetAsync(flag.p, 0, 4, s));
  hipLaunchKernelGGL(fill_f32_kernel, dim3(static_cast<unsigned>((2 * npad + 255) / 256)), dim3(256),
                     0, s, terms, 2 * npad, INFINITY);
  if (n > 0)
    hipLaunchKernelGGL(prep_rows_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       s, static_cast<const float*>(ix->rows), n, static_cast<int>(ix->ld),
                       static_cast<uint16_t*>(b), terms, terms + npad, flag.as<uint32_t>());
  SC_HIP(hipGetLastError());
  uint32_t bad = 0;
  SC_HIP(hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  ix->rows_bf16 = b;
  ix->norms2 = terms;  // [0, npad): y2m, [npad, 2 npad): nyu
  ix->ldb = bad ? 0 : static_cast<size_t>(npad);
  *bad_out = bad != 0;
  return SC_OK;
}

}  // namespace

int knn_mixed_device(const sc_index* cix, const void* d_queries, size_t q_, size_t k_,
                     int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t stream,
                     sc_stats* stats) {
  sc_index* ix = const_cast<sc_index*>(cix);
  const int64_t q = static_cast<int64_t>(q_);
  const int k = static_cast<int>(k_);
  bool bad = false;
  if (ix->ld > 16 * kMaxKS || k > kMaxK) bad = true;
  if (!bad)
    if (int rc = ensure_bf16(ix, stream, &bad)) return rc;
  if (bad)  // outside the bound's domain: the exact form serves the call
    return knn_exact_device(ix, d_queries, q_, k_, SC_EUCLIDIAN, 2, d_idx, d_dist, d_count,
                            stream, stats);
  const int64_t n = static_cast<int64_t>(ix->n);
  const int ld = static_cast<int>(ix->ld);
  const int64_t npad = static_cast<int64_t>(ix->ldb);
  const float* y2m = ix->norms2;
  const float* nyu = ix->norms2 + npad;

  // plan: ~256 workgroups (one per CU) over query tiles x row segments
  const int64_t nqt = (q + kQPB - 1) / kQPB;
  const int64_t tiles = (n + kRows - 1) / kRows;
  int64_t nseg = std::max<int64_t>(1, (256 + nqt - 1) / nqt);
  nseg = std::min<int64_t>(nseg, std::max<int64_t>(1, tiles));
  const int64_t seg_tiles = (tiles + nseg - 1) / nseg;
  nseg = std::max<int64_t>(1, (tiles + seg_tiles - 1) / seg_tiles);
  const int64_t seg_len = seg_tiles * kRows;
  const double Lseg = static_cast<double>(std::min<int64_t>(seg_len, std::max<int64_t>(n, 1)));
  int64_t cap = static_cast<int64_t>(4.0 * k * (1.0 + std::log(std::max(1.0, Lseg / k)))) + 128;
  cap = std::min<int64_t>(cap, seg_len);

  DevBuf qb, x2, cx, cidx, cL, ccnt, hs, lst, flags;
  SC_HIP(qb.alloc(sizeof(uint16_t) * q * ld, stream));
  SC_HIP(x2.alloc(sizeof(float) * q, stream));
  SC_HIP(cx.alloc(sizeof(float) * q, stream));
  SC_HIP(cidx.alloc(sizeof(uint32_t) * q * nseg * cap, stream));
  SC_HIP(cL.alloc(sizeof(float) * q * nseg * cap, stream));
  SC_HIP(ccnt.alloc(sizeof(uint32_t) * q * nseg, stream));
  SC_HIP(hs.alloc(sizeof(double) * q * k, stream));
  SC_HIP(lst.alloc(sizeof(uint32_t) * q, stream));
  SC_HIP(flags.alloc(64, stream));
  SC_HIP(hipMemsetAsync(flags.p, 0, 64, stream));
  uint32_t* list_n = flags.as<uint32_t>();
  unsigned long long* counters = reinterpret_cast<unsigned long long*>(flags.as<char>() + 8);
  const float* qf = static_cast<const float*>(d_queries);

  hipLaunchKernelGGL(prep_queries_kernel, dim3(static_cast<unsigned>((q + 255) / 256)), dim3(256),
                     0, stream, qf, q, ld, qb.as<uint16_t>(), x2.as<float>(), cx.as<float>(),
                     lst.as<uint32_t>(), list_n);
  SC_HIP(hipGetLastError());
  {
    ProfScope ps("knn_mixed_gemm", stream);
    dim3 grid(static_cast<unsigned>(nqt), static_cast<unsigned>(nseg));
    hipLaunchKernelGGL(mixed_scan_kernel, grid, dim3(256), 0, stream,
                       static_cast<const uint16_t*>(ix->rows_bf16), y2m, nyu, n, ld,
                       qb.as<uint16_t>(), x2.as<float>(), cx.as<float>(), q, k, seg_len,
                       static_cast<int>(nseg), cidx.as<uint32_t>(), cL.as<float>(),
                       ccnt.as<uint32_t>(), static_cast<int>(cap));
    SC_HIP(hipGetLastError());
  }
  {
    ProfScope ps("knn_mixed_rescore", stream);
    hipLaunchKernelGGL(mixed_replay_kernel, dim3(static_cast<unsigned>((q + 63) / 64)), dim3(64), 0,
                       stream, static_cast<const float*>(ix->rows), qf, ld, q, k,
                       static_cast<int>(nseg), static_cast<int>(cap), cidx.as<uint32_t>(),
                       cL.as<float>(), ccnt.as<uint32_t>(), x2.as<float>(), hs.as<double>(), d_idx,
                       d_dist, d_count, lst.as<uint32_t>(), list_n, counters);
    SC_HIP(hipGetLastError());
  }
  if (int rc = launch_serial_exact_f32(ix, qf, k, lst.as<uint32_t>(), list_n, d_idx, d_dist,
                                       d_count, stream))
    return rc;
  if (stats) {
    uint64_t h[3] = {0, 0, 0};
    uint32_t ln = 0;
    SC_HIP(hipMemcpyAsync(&ln, list_n, 4, hipMemcpyDeviceToHost, stream));
    SC_HIP(hipMemcpyAsync(h, counters, 16, hipMemcpyDeviceToHost, stream));
    SC_HIP(hipStreamSynchronize(stream));
    stats->queries += static_cast<uint64_t>(q);
    stats->candidates += h[0];
    stats->shortlist_rescored += h[1];
    stats->overflow_queries += ln;
    stats->uncertified += ln;
  }
  return SC_OK;
}

}  // namespace sc
