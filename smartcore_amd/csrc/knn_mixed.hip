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
// rounding inside s and every other FP32 rounding are covered by a relative
// margin of 2^-18 (x2 + y2).  Hence
//   L = (1 - 2^-18)(x2 + y2) - 2C - 2 c1 ||x|| ||y||  <=  s  <=
//   U = (1 + 2^-16)(x2 + y2) - 2C + 2 c1 ||x|| ||y||.
//
// The MFMA computes L - x2m directly: every BF16 row carries one extra
// 16-wide K step [y2 as three BF16 terms, nyb, 0...] and every BF16 query
// [-2x | 1, 1, 1, -cxb, 0...] (scaling by -2 is exact), with nyb >= ||y|| and
// cxb >= 2 c1 ||x|| rounded up to BF16 so their product over-covers the
// bound.  That step is issued last, after the cross terms, so its rounding
// is relative to the final magnitude and inside the 2^-18 margin.
//
// Pass 1 (scan, MFMA).  A workgroup of four waves owns 512 queries (128 per
// wave, held as BF16 B-fragments in registers) and one contiguous segment
// of rows, streamed 32 rows at a time (prefetched one tile ahead) through
// v_mfma_f32_32x32x16_bf16.  Workgroups are laid out so that the 32 CUs of
// an XCD stream the same segment and share its tiles in their L2.  Per
// query the wave keeps the k smallest U seen in the segment; every row whose
// L is below that running k-th U is appended (index, L) to the query's list
// for the segment, in index order.  A row that enters the reference heap has
// s below the running k-th smallest s of all earlier rows, which is at most
// the running k-th U of the segment, so the lists are a superset of the heap
// entrants.
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
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kQPW = 128;                 // queries per wave (4 column blocks)
constexpr int kQPB = kWaves * kQPW;       // queries per workgroup
constexpr int kRows = 32;                 // rows per tile
constexpr int kMaxKS = 8;                 // K steps of 16 -> dimension <= 128
constexpr int kMaxK = 32;                 // largest k served by this form
constexpr int kXcds = 8;                  // MI355X: 8 XCDs, round-robin dispatch
constexpr double kC1 = 0.0078125 + 1.52587890625e-05 + 128.0 * 2.384185791015625e-07;
constexpr double kLoMargin = 3.814697265625e-06;   // 2^-18
constexpr double kHiMargin = 1.52587890625e-05;    // 2^-16
constexpr double kNormLimit = 1e30;                // squared norms beyond this use FP64

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Smallest BF16 >= f for finite f >= 0.
__device__ __forceinline__ uint16_t bf16_ru_pos(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t t = u >> 16;
  return static_cast<uint16_t>((u & 0xffffu) ? t + 1 : t);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}

// Per row: BF16 copy with the extra K step, y2m = RD((1 - 2^-18) y2) and
// nyb = the BF16 value >= ||y||.  A row whose values are not finite or whose
// squared norm reaches kNormLimit marks `bad` (the index then uses FP64).
__global__ void prep_rows_kernel(const float* __restrict__ rows, int64_t n, int ld, int ldb,
                                 uint16_t* __restrict__ out_b, float* __restrict__ y2m,
                                 float* __restrict__ nyb, uint32_t* __restrict__ bad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rows + i * ld;
  uint16_t* o = out_b + i * ldb;
  double s = 0.0;
  bool ok = true;
  for (int j = 0; j < ld; ++j) {
    const float v = r[j];
    ok = ok && isfinite(v);
    s += static_cast<double>(v) * static_cast<double>(v);
    o[j] = bf16_rne(v);
  }
  for (int j = ld; j < ldb; ++j) o[j] = 0;
  if (!ok || !(s < kNormLimit)) {
    atomicOr(bad, 1u);
    y2m[i] = INFINITY;
    nyb[i] = INFINITY;
    return;
  }
  const float y2 = __double2float_rd(s * (1.0 - kLoMargin));
  const uint16_t h = bf16_rne(y2);
  const float r1 = y2 - bf16_to_f32(h);
  const uint16_t m = bf16_rne(r1);
  const float r2 = r1 - bf16_to_f32(m);
  const uint16_t l = bf16_rne(r2);
  const uint16_t nb = bf16_ru_pos(__double2float_ru(sqrt(s)));
  o[ld + 0] = h;
  o[ld + 1] = m;
  o[ld + 2] = l;
  o[ld + 3] = nb;
  y2m[i] = y2;
  nyb[i] = bf16_to_f32(nb);
}

__global__ void fill_f32_kernel(float* p, int64_t n, float v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per query: BF16 [-2x | 1, 1, 1, -cxb], x2m and cxb (the BF16 value
// >= 2 c1 ||x||); a query that is not finite or too large is listed for the
// exact kernel.
__global__ void prep_queries_kernel(const float* __restrict__ qs, int64_t q, int ld, int ldb,
                                    uint16_t* __restrict__ out_b, float* __restrict__ x2m,
                                    float* __restrict__ cxb, uint32_t* __restrict__ list,
                                    uint32_t* __restrict__ list_n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= q) return;
  const float* r = qs + i * ld;
  uint16_t* o = out_b + i * ldb;
  double s = 0.0;
  bool ok = true;
  for (int j = 0; j < ld; ++j) {
    const float v = r[j];
    ok = ok && isfinite(v);
    s += static_cast<double>(v) * static_cast<double>(v);
    o[j] = bf16_rne(-2.0f * v);
  }
  for (int j = ld; j < ldb; ++j) o[j] = 0;
  if (!ok || !(s < kNormLimit)) {
    // never a candidate in pass 1; the exact kernel takes it
    x2m[i] = INFINITY;
    cxb[i] = 0.0f;
    const uint32_t slot = atomicAdd(list_n, 1u);
    list[slot] = static_cast<uint32_t>(i);
    return;
  }
  const uint16_t cb = bf16_ru_pos(__double2float_ru(2.0 * kC1 * sqrt(s) * (1.0 + 9.5367431640625e-07)));
  const uint16_t one = 0x3f80;  // 1.0
  o[ld + 0] = one;
  o[ld + 1] = one;
  o[ld + 2] = one;
  o[ld + 3] = static_cast<uint16_t>(cb | 0x8000u);  // -cxb
  x2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  cxb[i] = bf16_to_f32(cb);
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(a, fminf(b, c)); }

template <int KS>
struct ScanCtx {
  static constexpr int KT = KS + 1;  // cross-term steps + the norm step
  static constexpr int LDB = 16 * KT;
};

template <int KS>
__global__ __launch_bounds__(256, 1) void mixed_scan_kernel(
    const uint16_t* __restrict__ rows_b, const float* __restrict__ y2m,
    const float* __restrict__ nyb, int64_t n, const uint16_t* __restrict__ q_b,
    const float* __restrict__ x2m, const float* __restrict__ cxbq, int64_t q, int k,
    int64_t seg_len, int nseg, int64_t nqt, uint32_t* __restrict__ cand_idx,
    float* __restrict__ cand_L, uint32_t* __restrict__ cand_count, int cap) {
  constexpr int KT = ScanCtx<KS>::KT;
  constexpr int LDB = ScanCtx<KS>::LDB;
  // XCD-aware layout: block b runs on XCD b % 8; the XCDs with the same
  // b % nseg stream the same segment with different query tiles.
  const int64_t bid = blockIdx.x;
  const int xcd = static_cast<int>(bid % kXcds);
  const int64_t slot = bid / kXcds;
  const int groups = kXcds / nseg;
  const int seg = xcd % nseg;
  const int64_t qt = slot * groups + xcd / nseg;
  if (qt >= nqt) return;  // whole workgroup leaves together

  __shared__ float topu[kQPB][kMaxK];  // per query: k smallest U, ascending
  __shared__ uint32_t ccount[kQPB];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 31;   // query column inside a 32-wide block / row of A
  const int half = lane >> 5;  // K half / row quad
  const int64_t qw0 = qt * kQPB + wave * kQPW;
  const int64_t r_begin = static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);

  for (int e = threadIdx.x; e < kQPB * kMaxK; e += 256) (&topu[0][0])[e] = INFINITY;
  for (int e = threadIdx.x; e < kQPB; e += 256) ccount[e] = 0u;
  __syncthreads();

  bf16x8 bq[4][KT];
  float thr[4], x2v[4];
  bool qvalid[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int64_t qi = qw0 + 32 * cb + col;
    // queries past the batch, and queries prep listed for the exact kernel
    // (x2m = inf), never produce candidates
    qvalid[cb] = qi < q && x2m[qi < q ? qi : 0] < INFINITY;
    const int64_t qr = qi < q ? qi : 0;
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) bq[cb][ks] = load_frag(q_b + qr * LDB + 16 * ks + 8 * half);
    x2v[cb] = qvalid[cb] ? x2m[qr] : INFINITY;
    thr[cb] = INFINITY;
  }
  const float cxq[4] = {qw0 + col < q ? cxbq[qw0 + col] : 0.0f,
                        qw0 + 32 + col < q ? cxbq[qw0 + 32 + col] : 0.0f,
                        qw0 + 64 + col < q ? cxbq[qw0 + 64 + col] : 0.0f,
                        qw0 + 96 + col < q ? cxbq[qw0 + 96 + col] : 0.0f};

  auto load_tile = [&](bf16x8 (&a)[KT], int64_t t0) {
    const int64_t arow = min<int64_t>(t0 + col, n - 1);
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) a[ks] = load_frag(rows_b + arow * LDB + 16 * ks + 8 * half);
  };

  auto process_tile = [&](const bf16x8 (&a)[KT], int64_t t0) {
    // four independent accumulation chains, interleaved so consecutive
    // MFMAs never depend on each other; the norm step (ks = KS) goes last
    f32x16 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
      acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], bq[cb][0], f32x16{}, 0, 0, 0);
#pragma unroll
    for (int ks = 1; ks < KT; ++ks)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
        acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], bq[cb][ks], acc[cb], 0, 0, 0);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      // acc = L - x2m for query (32 cb + col) and this lane's 16 rows
      float m = min3f(acc[cb][0], acc[cb][1], acc[cb][2]);
      m = min3f(m, acc[cb][3], acc[cb][4]);
      m = min3f(m, acc[cb][5], acc[cb][6]);
      m = min3f(m, acc[cb][7], acc[cb][8]);
      m = min3f(m, acc[cb][9], acc[cb][10]);
      m = min3f(m, acc[cb][11], acc[cb][12]);
      m = min3f(m, acc[cb][13], acc[cb][14]);
      m = fminf(m, acc[cb][15]);
      const bool any = qvalid[cb] && m < thr[cb];
      if (__ballot(any) == 0) continue;
      // Rare path: in row order, each lane holding row r for its query
      // inserts U into the query's running top-k and appends (row, L).
      const int ql = wave * kQPW + 32 * cb + col;
      const int64_t qi = qw0 + 32 * cb + col;
      for (int r = 0; r < kRows; ++r) {
        const int h = (r >> 2) & 1;
        const int v = (r >> 3) * 4 + (r & 3);
        const int64_t row = t0 + r;
        if (half == h && qvalid[cb] && row < r_end && acc[cb][v] < thr[cb]) {
          const double val = acc[cb][v];
          const double e = static_cast<double>(cxq[cb]) * static_cast<double>(nyb[row]);
          const double L = static_cast<double>(x2v[cb]) + val;
          const double U = L + 2.0 * e +
                           (kHiMargin + 2.0 * kLoMargin) *
                               (static_cast<double>(x2v[cb]) + static_cast<double>(y2m[row]));
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
            const size_t sl = (static_cast<size_t>(qi) * nseg + seg) * cap + c;
            cand_idx[sl] = static_cast<uint32_t>(row);
            cand_L[sl] = __double2float_rd(L);
          }
          ccount[ql] = c + 1;
        }
      }
      // T - x2m rounded up (the f64 difference of two f32 values is exact)
      thr[cb] = __double2float_ru(static_cast<double>(topu[ql][k - 1]) -
                                  static_cast<double>(x2v[cb]));
    }
  };

  bf16x8 a0[KT], a1[KT];
  int64_t t0 = r_begin;
  if (t0 < r_end) load_tile(a0, t0);
  while (t0 < r_end) {
    if (t0 + kRows < r_end) load_tile(a1, t0 + kRows);
    process_tile(a0, t0);
    t0 += kRows;
    if (t0 >= r_end) break;
    if (t0 + kRows < r_end) load_tile(a0, t0 + kRows);
    process_tile(a1, t0);
    t0 += kRows;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kQPB; e += 256) {
    const int64_t qi = qt * kQPB + e;
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

// Builds the BF16 copy and row terms of an f32 index once.  ldb (the BF16
// row stride, ld + 16) is recorded in the index; 0 marks an index outside
// the bound's domain.
int ensure_bf16(sc_index* ix, hipStream_t s, bool* bad_out, int64_t* npad_out) {
  std::lock_guard<std::mutex> lk(*static_cast<std::mutex*>(ix->mu));
  const int64_t n = static_cast<int64_t>(ix->n);
  const int64_t npad = (n + kRows - 1) / kRows * kRows + kRows;
  *npad_out = npad;
  if (ix->rows_bf16) {
    *bad_out = ix->ldb == 0;
    return SC_OK;
  }
  const int ldb = static_cast<int>(ix->ld) + 16;
  void* b = nullptr;
  float* terms = nullptr;
  SC_HIP(hipMalloc(&b, std::max<size_t>(1, n) * ldb * sizeof(uint16_t)));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&terms), sizeof(float) * 2 * npad));
  DevBuf flag;
  SC_HIP(flag.alloc(4, s));
  SC_HIP(hipMemsetAsync(flag.p, 0, 4, s));
  hipLaunchKernelGGL(fill_f32_kernel, dim3(static_cast<unsigned>((2 * npad + 255) / 256)), dim3(256),
                     0, s, terms, 2 * npad, INFINITY);
  if (n > 0)
    hipLaunchKernelGGL(prep_rows_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       s, static_cast<const float*>(ix->rows), n, static_cast<int>(ix->ld), ldb,
                       static_cast<uint16_t*>(b), terms, terms + npad, flag.as<uint32_t>());
  SC_HIP(hipGetLastError());
  uint32_t bad = 0;
  SC_HIP(hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  ix->rows_bf16 = b;
  ix->norms2 = terms;  // [0, npad): y2m, [npad, 2 npad): nyb
  ix->ldb = bad ? 0 : static_cast<size_t>(ldb);
  *bad_out = bad != 0;
  return SC_OK;
}

template <int KS>
void launch_scan(dim3 grid, hipStream_t stream, const uint16_t* rows_b, const float* y2m,
                 const float* nyb, int64_t n, const uint16_t* q_b, const float* x2m,
                 const float* cxb, int64_t q, int k, int64_t seg_len, int nseg, int64_t nqt,
                 uint32_t* cidx, float* cL, uint32_t* ccnt, int cap) {
  hipLaunchKernelGGL(mixed_scan_kernel<KS>, grid, dim3(256), 0, stream, rows_b, y2m, nyb, n, q_b,
                     x2m, cxb, q, k, seg_len, nseg, nqt, cidx, cL, ccnt, cap);
}

}  // namespace

int knn_mixed_device(const sc_index* cix, const void* d_queries, size_t q_, size_t k_,
                     int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t stream,
                     sc_stats* stats) {
  sc_index* ix = const_cast<sc_index*>(cix);
  const int64_t q = static_cast<int64_t>(q_);
  const int k = static_cast<int>(k_);
  bool bad = false;
  int64_t npad = 0;
  if (ix->ld > 16 * kMaxKS || k > kMaxK) bad = true;
  if (!bad)
    if (int rc = ensure_bf16(ix, stream, &bad, &npad)) return rc;
  if (bad)  // outside the bound's domain: the exact form serves the call
    return knn_exact_device(ix, d_queries, q_, k_, SC_EUCLIDIAN, 2, d_idx, d_dist, d_count,
                            stream, stats);
  const int64_t n = static_cast<int64_t>(ix->n);
  const int ld = static_cast<int>(ix->ld);
  const int ldb = static_cast<int>(ix->ldb);
  const float* y2m = ix->norms2;
  const float* nyb = ix->norms2 + npad;

  // plan: ~256 workgroups (one per CU); segments in {1, 2, 4, 8} so that
  // whole XCDs share a segment
  const int64_t nqt = (q + kQPB - 1) / kQPB;
  const int64_t tiles = std::max<int64_t>(1, (n + kRows - 1) / kRows);
  int nseg = 1;
  while (nseg < kXcds && nqt * nseg < 256 && nseg * 2 <= tiles) nseg *= 2;
  const int64_t seg_tiles = (tiles + nseg - 1) / nseg;
  const int64_t seg_len = seg_tiles * kRows;
  const int groups = kXcds / nseg;
  const int64_t slots = (nqt + groups - 1) / groups;
  const double Lseg = static_cast<double>(std::min<int64_t>(seg_len, std::max<int64_t>(n, 1)));
  int64_t cap = static_cast<int64_t>(4.0 * k * (1.0 + std::log(std::max(1.0, Lseg / k)))) + 128;
  cap = std::min<int64_t>(cap, seg_len);

  DevBuf qb, x2, cx, cidx, cL, ccnt, hs, lst, flags;
  SC_HIP(qb.alloc(sizeof(uint16_t) * q * ldb, stream));
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
                     0, stream, qf, q, ld, ldb, qb.as<uint16_t>(), x2.as<float>(), cx.as<float>(),
                     lst.as<uint32_t>(), list_n);
  SC_HIP(hipGetLastError());
  {
    ProfScope ps("knn_mixed_gemm", stream);
    const dim3 grid(static_cast<unsigned>(slots * kXcds));
    const uint16_t* rb = static_cast<const uint16_t*>(ix->rows_bf16);
    switch (ld / 16) {
#define SC_SCAN_CASE(KS)                                                                        \
  case KS:                                                                                     \
    launch_scan<KS>(grid, stream, rb, y2m, nyb, n, qb.as<uint16_t>(), x2.as<float>(),          \
                    cx.as<float>(), q, k, seg_len, nseg, nqt, cidx.as<uint32_t>(), cL.as<float>(), \
                    ccnt.as<uint32_t>(), static_cast<int>(cap));                               \
    break;
      SC_SCAN_CASE(1)
      SC_SCAN_CASE(2)
      SC_SCAN_CASE(3)
      SC_SCAN_CASE(4)
      SC_SCAN_CASE(5)
      SC_SCAN_CASE(6)
      SC_SCAN_CASE(7)
      SC_SCAN_CASE(8)
#undef SC_SCAN_CASE
      default:
        return fail(SC_INVALID_STATE, "unsupported row stride for the mixed form");
    }
    SC_HIP(hipGetLastError());
  }
  {
    ProfScope ps("knn_mixed_rescore", stream);
    hipLaunchKernelGGL(mixed_replay_kernel, dim3(static_cast<unsigned>((q + 63) / 64)), dim3(64), 0,
                       stream, static_cast<const float*>(ix->rows), qf, ld, q, k, nseg,
                       static_cast<int>(cap), cidx.as<uint32_t>(), cL.as<float>(),
                       ccnt.as<uint32_t>(), x2.as<float>(), hs.as<double>(), d_idx, d_dist, d_count,
                       lst.as<uint32_t>(), list_n, counters);
    SC_HIP(hipGetLastError());
  }
  if (int rc = launch_serial_exact_f32(ix, qf, k, lst.as<uint32_t>(), list_n, d_idx, d_dist,
                                       d_count, stream))
    return rc;
  if (stats) {
    uint64_t h[2] = {0, 0};
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
