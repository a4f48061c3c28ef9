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
// Pass 1 (scan, MFMA).  A workgroup of eight waves owns 512 queries (64 per
// wave, held as BF16 B-fragments of -2x in registers; scaling by -2 is
// exact) and one contiguous segment of rows, streamed 32 rows at a time
// through three padded LDS tile slots that all eight waves read (each tile
// is loaded into registers two tiles ahead).  The loop is software
// pipelined: a wave issues the MFMAs of tile t + 1 before the epilogue of
// tile t, so the matrix pipe runs under the epilogue, the rare candidate
// path and the tile store.  v_mfma_f32_32x32x16_bf16 gives
// acc = -2C; the epilogue forms L - x2m = y2m - cxb nyb + acc with one FMA
// and one add (y2m <= (1 - 2^-18) y2, nyb >= ||y||, cxb >= 2 c1 ||x||, all
// rounded the safe way; the two roundings are inside the 2^-18 margin).
// Workgroups are laid out so that the CUs of an XCD stream the same segment
// and share its tiles in their L2.  Per query the wave keeps the k smallest
// U seen in the segment; every row whose L is below that running k-th U is
// appended (index, L) to the query's list for the segment, in index order.
// A row that enters the reference heap has s below the running k-th
// smallest s of all earlier rows, which is at most the running k-th U of the
// segment, so the lists are a superset of the heap entrants.
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

constexpr int kWaves = 8;                 // waves per workgroup (two per SIMD)
constexpr int kQPW = 64;                  // queries per wave (2 column blocks)
constexpr int kQPB = kWaves * kQPW;       // queries per workgroup
constexpr int kNT = 64 * kWaves;          // threads per workgroup
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

// Per row: BF16 copy, y2m = RD((1 - 2^-18) y2) and nyb = RU(||y||) (with a
// relative allowance for the f64 sum behind y2).  A row whose values are not
// finite or whose squared norm reaches kNormLimit marks `bad` (the index
// then uses FP64).
__global__ void prep_rows_kernel(const float* __restrict__ rows, int64_t n, int ld,
                                 uint16_t* __restrict__ out_b, float* __restrict__ y2m,
                                 float* __restrict__ nyb, uint32_t* __restrict__ bad) {
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
    o[j] = bf16_rne(v);
  }
  if (!ok || !(s < kNormLimit)) {
    atomicOr(bad, 1u);
    y2m[i] = INFINITY;
    nyb[i] = 0.0f;
    return;
  }
  y2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  nyb[i] = __double2float_ru(sqrt(s) * (1.0 + 9.094947017729282e-13));
}

// y2m of padding rows is +inf (their L is +inf), nyb is 0.
__global__ void fill_terms_kernel(float* y2m, float* nyb, int64_t from, int64_t npad) {
  const int64_t i = from + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < npad) {
    y2m[i] = INFINITY;
    nyb[i] = 0.0f;
  }
}

// Per query: BF16 of -2x, x2m = RD((1 - 2^-18) x2) and cxb = RU(2 c1 ||x||);
// a query that is not finite or too large is listed for the exact kernel.
__global__ void prep_queries_kernel(const float* __restrict__ qs, int64_t q, int ld,
                                    uint16_t* __restrict__ out_b, float* __restrict__ x2m,
                                    float* __restrict__ cxb, uint32_t* __restrict__ list,
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
    o[j] = bf16_rne(-2.0f * v);
  }
  if (!ok || !(s < kNormLimit)) {
    // never a candidate in pass 1; the exact kernel takes it
    x2m[i] = INFINITY;
    cxb[i] = 0.0f;
    const uint32_t slot = atomicAdd(list_n, 1u);
    list[slot] = static_cast<uint32_t>(i);
    return;
  }
  x2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  cxb[i] = __double2float_ru(2.0 * kC1 * sqrt(s) * (1.0 + 9.5367431640625e-07));
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

template <int KS>
struct Scan {
  static constexpr int LD = 16 * KS;                    // row length (bf16)
  static constexpr int RS = LD + 8;                     // LDS row stride: 16 B of padding
  static constexpr int VEC = kRows * LD / 8;            // 16-byte pieces of a tile
  static constexpr int SLOT = kRows * RS * 2 + 2 * kRows * 4;  // rows | y2m | nyb (bytes)
  static_assert(VEC <= kNT, "one piece per thread");
};

// Row stride of the per-query top-k: 16 (one aligned register block) up to
// k = 16, else k.
__host__ __device__ inline int topk_stride(int k) { return k <= 16 ? 16 : k; }

// Dynamic LDS of the scan: three tile slots | topu[kQPB][stride] |
// ccount[kQPB] | trans[kWaves][64][16].
template <int KS>
size_t scan_lds_bytes(int k) {
  return 3 * static_cast<size_t>(Scan<KS>::SLOT) + sizeof(float) * kQPB * topk_stride(k) +
         sizeof(uint32_t) * kQPB + sizeof(float) * kWaves * 64 * 16;
}

template <int KS>
__global__ __launch_bounds__(kNT, 2) void mixed_scan_kernel(
    const uint16_t* __restrict__ rows_b, const float* __restrict__ y2m,
    const float* __restrict__ nyb, int64_t n, const uint16_t* __restrict__ q_b,
    const float* __restrict__ x2m, const float* __restrict__ cxbq, int64_t q, int k,
    int64_t seg_len, int nseg, int64_t nqt, uint32_t* __restrict__ cand_idx,
    float* __restrict__ cand_L, uint32_t* __restrict__ cand_count, int cap) {
  using S = Scan<KS>;
  extern __shared__ __align__(16) unsigned char smem[];
  // XCD-aware layout: block b runs on XCD b % 8.  Up to 8 segments, the
  // XCDs with the same b % nseg stream the same segment with different
  // query tiles; beyond 8, each XCD takes nseg / 8 segments of its own.
  const int64_t bid = blockIdx.x;
  const int xcd = static_cast<int>(bid % kXcds);
  const int64_t slot_i = bid / kXcds;
  int seg;
  int64_t qt;
  if (nseg <= kXcds) {
    const int groups = kXcds / nseg;
    seg = xcd % nseg;
    qt = slot_i * groups + xcd / nseg;
  } else {
    const int m = nseg / kXcds;
    seg = static_cast<int>(slot_i % m) * kXcds + xcd;
    qt = slot_i / m;
  }
  if (qt >= nqt) return;  // whole workgroup leaves together

  const int kst = topk_stride(k);
  float* topu = reinterpret_cast<float*>(smem + 3 * S::SLOT);  // per query: k smallest U
  uint32_t* ccount = reinterpret_cast<uint32_t*>(topu + kQPB * kst);
  float* trans = reinterpret_cast<float*>(ccount + kQPB);      // slow-path transpose
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int col = lane & 31;   // query column inside a 32-wide block / row of A
  const int half = lane >> 5;  // K half / row quad
  const int64_t qw0 = qt * kQPB + wave * kQPW;
  const int64_t r_begin = static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);
  const int64_t ntiles = r_end > r_begin ? (r_end - r_begin + kRows - 1) / kRows : 0;

  for (int e = tid; e < kQPB * kst; e += kNT) topu[e] = INFINITY;
  for (int e = tid; e < kQPB; e += kNT) ccount[e] = 0u;

  bf16x8 bq[2][KS];
  float thr[2], x2v[2], cxq[2];
  bool qvalid[2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int64_t qi = qw0 + 32 * cb + col;
    // queries past the batch, and queries prep listed for the exact kernel
    // (x2m = inf), never produce candidates
    qvalid[cb] = qi < q && x2m[qi < q ? qi : 0] < INFINITY;
    const int64_t qr = qi < q ? qi : 0;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) bq[cb][ks] = load_frag(q_b + qr * S::LD + 16 * ks + 8 * half);
    x2v[cb] = qvalid[cb] ? x2m[qr] : INFINITY;
    cxq[cb] = qi < q ? cxbq[qr] : 0.0f;
    thr[cb] = INFINITY;
  }

  // tile loads: thread tid < VEC carries 16 bytes of rows; threads 0..63
  // carry the tile's y2m (0..31) and nyb (32..63)
  const int pr_row = tid / (S::LD / 8), pr_c8 = tid % (S::LD / 8);
  auto gload = [&](int64_t t, u32x4& v, float& tv) {
    const int64_t t0 = r_begin + t * kRows;
    if (tid < S::VEC) {
      const int64_t row = min<int64_t>(t0 + pr_row, n - 1);
      v = *reinterpret_cast<const u32x4*>(rows_b + row * S::LD + 8 * pr_c8);
    }
    if (tid < 64) tv = (tid < 32 ? y2m : nyb)[t0 + (tid & 31)];
  };
  auto swrite = [&](int s, const u32x4& v, float tv) {
    unsigned char* base = smem + s * S::SLOT;
    if (tid < S::VEC)
      *reinterpret_cast<u32x4*>(base + 2 * (pr_row * S::RS + 8 * pr_c8)) = v;
    if (tid < 64) reinterpret_cast<float*>(base + 2 * kRows * S::RS)[tid] = tv;
  };

  // Software pipeline, one barrier per tile.  LDS holds three tiles: the
  // one whose MFMAs are issued (t + 1), the one whose epilogue runs (t; its
  // row terms are still needed) and the one being written (t + 2).  Each
  // wave issues the MFMAs of tile t + 1 interleaved with the VALU epilogue
  // of tile t, so the matrix pipe keeps working while the epilogue, the
  // rare candidate path and the tile store run.
  auto slot_ptr = [&](int64_t t) { return smem + static_cast<int>(t % 3) * S::SLOT; };
  auto mfma_tile = [&](int64_t t, f32x16 (&acc)[2]) {
    const uint16_t* A = reinterpret_cast<const uint16_t*>(slot_ptr(t));
    acc[0] = f32x16{};
    acc[1] = f32x16{};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const bf16x8 a = load_frag(A + col * S::RS + 16 * ks + 8 * half);
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[0][ks], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[1][ks], acc[1], 0, 0, 0);
    }
  };
  // fast epilogue of tile t: acc <- L - x2m; returns per column block
  // whether any lane of the wave has a row below its query's threshold
  auto epilogue = [&](int64_t t, f32x16 (&acc)[2], bool (&hit)[2]) {
    const unsigned char* base = slot_ptr(t);
    const float* ty2 = reinterpret_cast<const float*>(base + 2 * kRows * S::RS);
    const float* tny = ty2 + kRows;
    f32x4 yv[4], nv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      yv[g] = *reinterpret_cast<const f32x4*>(ty2 + 8 * g + 4 * half);
      nv[g] = *reinterpret_cast<const f32x4*>(tny + 8 * g + 4 * half);
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      float m = INFINITY;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = fmaf(-cxq[cb], nv[g][r], acc[cb][4 * g + r]) + yv[g][r];
          acc[cb][4 * g + r] = v;
          m = fminf(m, v);
        }
      hit[cb] = __ballot(qvalid[cb] && m < thr[cb]) != 0;
    }
  };
  // Rare path for column block cb of tile t (see the file comment).
  auto slow = [&](int64_t t, int cb, f32x16& accv) {
    const unsigned char* base = slot_ptr(t);
    const float* ty2 = reinterpret_cast<const float*>(base + 2 * kRows * S::RS);
    const float* tny = ty2 + kRows;
    const int64_t t0 = r_begin + t * kRows;
    uint32_t m16 = 0;
#pragma unroll
    for (int v = 0; v < 16; ++v) m16 |= static_cast<uint32_t>(qvalid[cb] && accv[v] < thr[cb]) << v;
    float* tw = trans + (wave * 64 + lane) * 16;
#pragma unroll
    for (int g = 0; g < 4; ++g)
      *reinterpret_cast<f32x4*>(tw + 4 * g) =
          f32x4{accv[4 * g], accv[4 * g + 1], accv[4 * g + 2], accv[4 * g + 3]};
    const uint32_t pm = __shfl_xor(m16, 32);
    uint32_t M = 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      M |= ((m16 >> (4 * g)) & 0xFu) << (8 * g);
      M |= ((pm >> (4 * g)) & 0xFu) << (8 * g + 4);
    }
    float tcur = thr[cb];
    if (half == 0 && M != 0) {
      const int ql = wave * kQPW + 32 * cb + col;
      const int64_t qi = qw0 + 32 * cb + col;
      float* tk = topu + ql * kst;
      uint32_t c = ccount[ql];
      const double xv = static_cast<double>(x2v[cb]);
      // k <= 16: the sorted top-k lives in registers while this tile's rows
      // are inserted (one LDS round trip instead of a dependent chain)
      const bool vec = k <= 16;
      float tv[16];
      if (vec) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const f32x4 q4 = *reinterpret_cast<const f32x4*>(tk + 4 * g);
          tv[4 * g] = q4[0];
          tv[4 * g + 1] = q4[1];
          tv[4 * g + 2] = q4[2];
          tv[4 * g + 3] = q4[3];
        }
      }
      while (M != 0) {
        const int r = __builtin_ctz(M);
        M &= M - 1;
        const int64_t row = t0 + r;
        if (row >= r_end) break;  // rows ascend: the rest are past the segment
        const int h = (r >> 2) & 1;
        const int v = (r >> 3) * 4 + (r & 3);
        const float val = trans[(wave * 64 + col + 32 * h) * 16 + v];
        if (!(val < tcur)) continue;  // the threshold fell earlier in this tile
        const double e = static_cast<double>(cxq[cb]) * static_cast<double>(tny[r]);
        const double L = xv + static_cast<double>(val);
        const double U = L + 2.0 * e + (kHiMargin + 2.0 * kLoMargin) * (xv + static_cast<double>(ty2[r]));
        const float uf = __double2float_ru(U);
        float kth;
        if (vec) {
          // insert uf into the ascending tv[0..k): slot j keeps its value if
          // it is <= uf, else takes uf or its left neighbour
          float nv[16];
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            const float left = j ? tv[j - 1] : -INFINITY;
            nv[j] = j >= k ? INFINITY : (tv[j] <= uf ? tv[j] : (left <= uf ? uf : left));
          }
          kth = INFINITY;
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            tv[j] = nv[j];
            if (j == k - 1) kth = nv[j];
          }
        } else {
          if (uf < tk[k - 1]) {
            int pos = k - 1;
            while (pos > 0 && tk[pos - 1] > uf) {
              tk[pos] = tk[pos - 1];
              --pos;
            }
            tk[pos] = uf;
          }
          kth = tk[k - 1];
        }
        if (c < static_cast<uint32_t>(cap)) {
          const size_t sl = (static_cast<size_t>(qi) * nseg + seg) * cap + c;
          cand_idx[sl] = static_cast<uint32_t>(row);
          cand_L[sl] = __double2float_rd(L);
        }
        ++c;
        // T - x2m rounded up (the f64 difference of two f32 values is exact)
        tcur = __double2float_ru(static_cast<double>(kth) - xv);
      }
      if (vec) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
          *reinterpret_cast<f32x4*>(tk + 4 * g) =
              f32x4{tv[4 * g], tv[4 * g + 1], tv[4 * g + 2], tv[4 * g + 3]};
      }
      ccount[ql] = c;
    }
    // both lanes of the query take the half-0 lane's threshold
    const float tn = __shfl_xor(tcur, 32);
    thr[cb] = half == 0 ? tcur : tn;
  };
  // one pipeline step: MFMAs of tile t + 1 into nxt, epilogue of tile t in
  // cur, then the rare path, the store of tile t + 2 and the barrier
  auto step = [&](int64_t t, f32x16 (&cur)[2], f32x16 (&nxt)[2], u32x4& rv, float& tv) {
    bool hit[2];
    // tile t + 1 may be past the segment on the last step: its slot then
    // holds stale rows whose results are never read
    mfma_tile(t + 1, nxt);
    epilogue(t, cur, hit);
    if (hit[0]) slow(t, 0, cur[0]);
    if (hit[1]) slow(t, 1, cur[1]);
    if (t + 2 < ntiles) {
      swrite(static_cast<int>((t + 2) % 3), rv, tv);
      if (t + 4 < ntiles) gload(t + 4, rv, tv);
    }
    __syncthreads();
  };

  u32x4 ra = {0u, 0u, 0u, 0u}, rb = {0u, 0u, 0u, 0u};
  float ta = 0.0f, tb = 0.0f;
  f32x16 accA[2], accB[2];
  if (ntiles > 0) gload(0, ra, ta);
  if (ntiles > 1) gload(1, rb, tb);
  if (ntiles > 0) swrite(0, ra, ta);
  if (ntiles > 1) swrite(1, rb, tb);
  if (ntiles > 2) gload(2, ra, ta);
  if (ntiles > 3) gload(3, rb, tb);
  __syncthreads();
  if (ntiles > 0) mfma_tile(0, accA);
  for (int64_t t = 0; t < ntiles; t += 2) {
    // tile t: MFMAs in accA, tile t + 2 in ra; tile t + 3 in rb
    step(t, accA, accB, ra, ta);
    if (t + 1 >= ntiles) break;
    step(t + 1, accB, accA, rb, tb);
  }
  for (int e = tid; e < kQPB; e += kNT) {
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
// row stride, ld) is recorded in the index; 0 marks an index outside the
// bound's domain.
int ensure_bf16(sc_index* ix, hipStream_t s, bool* bad_out, int64_t* npad_out) {
  std::lock_guard<std::mutex> lk(*static_cast<std::mutex*>(ix->mu));
  const int64_t n = static_cast<int64_t>(ix->n);
  const int64_t npad = (n + kRows - 1) / kRows * kRows + kRows;
  *npad_out = npad;
  if (ix->rows_bf16) {
    *bad_out = ix->ldb == 0;
    return SC_OK;
  }
  const int ldb = static_cast<int>(ix->ld);
  void* b = nullptr;
  float* terms = nullptr;
  SC_HIP(hipMalloc(&b, std::max<size_t>(1, n) * ldb * sizeof(uint16_t)));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&terms), sizeof(float) * 2 * npad));
  DevBuf flag;
  SC_HIP(flag.alloc(4, s));
  SC_HIP(hipMemsetAsync(flag.p, 0, 4, s));
  hipLaunchKernelGGL(fill_terms_kernel, dim3(static_cast<unsigned>((npad - n + 255) / 256)), dim3(256),
                     0, s, terms, terms + npad, n, npad);
  if (n > 0)
    hipLaunchKernelGGL(prep_rows_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       s, static_cast<const float*>(ix->rows), n, static_cast<int>(ix->ld),
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
  const size_t lds = scan_lds_bytes<KS>(k);
  static const bool attr = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&mixed_scan_kernel<KS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(scan_lds_bytes<KS>(kMaxK)));
    return true;
  }();
  (void)attr;
  hipLaunchKernelGGL(mixed_scan_kernel<KS>, grid, dim3(kNT), lds, stream, rows_b, y2m, nyb, n, q_b,
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
  const float* y2m = ix->norms2;
  const float* nyb = ix->norms2 + npad;

  // plan: ~256 workgroups (one per CU).  Segments are a power of two: up to
  // 8 whole XCDs share a segment; small batches split the rows further so
  // that they still fill the chip.
  const int64_t nqt = (q + kQPB - 1) / kQPB;
  const int64_t tiles = std::max<int64_t>(1, (n + kRows - 1) / kRows);
  int nseg = 1;
  while (nseg < 64 && nqt * nseg < 256 && nseg * 2 <= tiles) nseg *= 2;
  const int64_t seg_tiles = (tiles + nseg - 1) / nseg;
  const int64_t seg_len = seg_tiles * kRows;
  const int64_t nblocks =
      nseg <= kXcds ? (nqt + kXcds / nseg - 1) / (kXcds / nseg) * kXcds : nqt * nseg;
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
    const dim3 grid(static_cast<unsigned>(nblocks));
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
