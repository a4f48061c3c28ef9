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
// through an eight-slot LDS tile ring that all waves read: tiles are loaded
// into registers four steps ahead and stored four at a time, with one
// barrier per four steps.  v_mfma_f32_32x32x16_bf16 gives acc = -2C.  In
// the running launch the epilogue forms L - x2m = y2m - cxb nyb + acc with
// one FMA and one add (y2m <= (1 - 2^-18) y2, nyb >= ||y||, cxb >= 2 c1
// ||x||, all rounded the safe way; the two roundings are inside the 2^-18
// margin); the fixed-threshold launch folds its bound into one more MFMA K
// step (at fixed_scan_kernel).  Workgroups are laid out so that the CUs of
// an XCD stream the same segment and share its tiles in their L2.  In the
// running launch each lane owns one query and keeps, in registers, the k
// smallest U seen so far; every row whose L is below that running k-th U is
// appended to the query's list for the segment, in index order.
// A row that enters the reference heap has s below the running k-th
// smallest s of all earlier rows, which is at most the running k-th U, so
// the lists are a superset of the heap entrants.
// Levels.  Small indexes run one such launch.  For indexes of at least
// kSeedMinRows rows the first launch covers only the first kPrefixRows
// rows; then fixed-threshold launches cover row ranges growing kLevelRatio
// times.  The candidates of each level get their exact s (Pass 2a) as the
// level ends; the next level tests every row against T, the k-th smallest
// s among the candidates of all earlier levels rounded up to f32, and
// appends the rows with L < T.  The candidates are a subset of the earlier
// rows, so T is at or above the k-th smallest s of all earlier rows, which
// is the heap's threshold when the level starts: these lists are also a
// superset of the entrants, and with no running threshold a lane writes
// its own rows without waiting on other lanes.  (The candidates also hold
// the k nearest earlier rows -- each had L < s below the threshold of its
// own level -- so T is the exact k-th distance, narrower than a k-th U by
// the width of the bound, and fewer rows pass.)
// Candidates are kept only if the per-row bound, which uses the stored
// norms of y - bf16(y) and x - bf16(x), also has L below the threshold.
// Pass 2 (replay, FP64).  Every candidate's exact s is computed in
// parallel (a workgroup per query); then one wavefront per query walks its
// lists in index order, 64 entries at a time, and applies the heap rule to
// the entries below the heap's top, one after another.  Rows that would not
// enter the reference heap are no-ops, so the final heap array equals the
// reference's.  Queries whose lists overflow, or whose rows are not finite
// / too large for the bound, are recomputed by the sequential exact kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

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
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

constexpr int kWaves = 8;                 // waves per workgroup (two per SIMD)
constexpr int kQPW = 64;                  // queries per wave (2 column blocks)
constexpr int kQPB = kWaves * kQPW;       // queries per workgroup
constexpr int kNT = 64 * kWaves;          // threads per workgroup
constexpr int kRows = 32;                 // rows per tile
constexpr int kMaxKS = 8;                 // K steps of 16 -> dimension <= 128
constexpr int kMaxK = 32;                 // largest k served by this form
constexpr int kXcds = 8;                  // MI355X: 8 XCDs, round-robin dispatch
constexpr int kTargetBlocks = 256;        // one workgroup per CU on 256 CUs
constexpr int64_t kPrefixRows = 16384;    // rows of the seeding launch
constexpr int64_t kSeedMinRows = 524288;  // indexes this large are seeded
constexpr int64_t kLevelRatio = 8;        // row-range growth of the fixed-threshold levels
constexpr int64_t kScanPadRows = 1024;    // padding rows past the index read by fixed launches
constexpr int kMaxLists = 64 + 2 * 64 * 8;  // running lists + fixed pairs, any plan
constexpr double kC1 = 0.0078125 + 1.52587890625e-05 + 128.0 * 2.384185791015625e-07;
constexpr double kLoMargin = 3.814697265625e-06;   // 2^-18
constexpr double kHiMargin = 1.52587890625e-05;    // 2^-16
constexpr double kNormLimit = 1e30;                // squared norms beyond this use FP64

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Per row: BF16 copy, y2m = RD((1 - 2^-18) y2), nyb = RU(||y||),
// nry = RU(||y - bf16(y)||) and nyh = RU(||y|| + ||y - bf16(y)||) >=
// ||bf16(y)|| (with a relative allowance for the f64 sums).
// A row whose values are not finite or whose squared norm reaches
// kNormLimit marks `bad` (the index then uses FP64).
__global__ void prep_rows_kernel(const float* __restrict__ rows, int64_t n, int ld,
                                 uint16_t* __restrict__ out_b, float* __restrict__ y2m,
                                 float* __restrict__ nyb, float* __restrict__ nry,
                                 float* __restrict__ nyh, uint32_t* __restrict__ bad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rows + i * ld;
  uint16_t* o = out_b + i * ld;
  double s = 0.0, sr = 0.0;
  bool ok = true;
  for (int j = 0; j < ld; ++j) {
    const float v = r[j];
    ok = ok && isfinite(v);
    s += static_cast<double>(v) * static_cast<double>(v);
    const uint16_t b = bf16_rne(v);
    o[j] = b;
    const double e = static_cast<double>(v) - static_cast<double>(__uint_as_float(static_cast<uint32_t>(b) << 16));
    sr += e * e;
  }
  if (!ok || !(s < kNormLimit)) {
    atomicOr(bad, 1u);
    y2m[i] = INFINITY;
    nyb[i] = 0.0f;
    nry[i] = 0.0f;
    nyh[i] = 0.0f;
    return;
  }
  y2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  nyb[i] = __double2float_ru(sqrt(s) * (1.0 + 9.094947017729282e-13));
  nry[i] = __double2float_ru(sqrt(sr) * (1.0 + 9.094947017729282e-13));
  nyh[i] = __double2float_ru((sqrt(s) + sqrt(sr)) * (1.0 + 9.094947017729282e-13));
}

// y2m of padding rows is +inf (their L is +inf), nyb, nry and nyh are 0.
__global__ void fill_terms_kernel(float* y2m, float* nyb, float* nry, float* nyh, int64_t from,
                                  int64_t npad) {
  const int64_t i = from + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < npad) {
    y2m[i] = INFINITY;
    nyb[i] = 0.0f;
    nry[i] = 0.0f;
    nyh[i] = 0.0f;
  }
}

// Per query: BF16 of -2x, x2m = RD((1 - 2^-18) x2), cxb = RU(2 c1 ||x||)
// and the terms of the tight bound (see slow() in the scan):
// qa1 = RU(2 ||x||), qa2 = RU(2 (||x - bf16(x)|| + 128 * 2^-22 ||bf16(x)||));
// a query that is not finite or too large is listed for the exact kernel.
__global__ void prep_queries_kernel(const float* __restrict__ qs, int64_t q, int ld,
                                    uint16_t* __restrict__ out_b, float* __restrict__ x2m,
                                    float* __restrict__ cxb, float* __restrict__ qa1,
                                    float* __restrict__ qa2, uint32_t* __restrict__ list,
                                    uint32_t* __restrict__ list_n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= q) return;
  const float* r = qs + i * ld;
  uint16_t* o = out_b + i * ld;
  double s = 0.0, sr = 0.0, sh = 0.0;
  bool ok = true;
  for (int j = 0; j < ld; ++j) {
    const float v = r[j];
    ok = ok && isfinite(v);
    s += static_cast<double>(v) * static_cast<double>(v);
    const uint16_t b = bf16_rne(-2.0f * v);
    o[j] = b;
    // bf16(-2v) = -2 bf16(v) exactly
    const double h = -0.5 * static_cast<double>(__uint_as_float(static_cast<uint32_t>(b) << 16));
    const double e = static_cast<double>(v) - h;
    sr += e * e;
    sh += h * h;
  }
  if (!ok || !(s < kNormLimit)) {
    // never a candidate in pass 1; the exact kernel takes it
    x2m[i] = INFINITY;
    cxb[i] = 0.0f;
    qa1[i] = 0.0f;
    qa2[i] = 0.0f;
    const uint32_t slot = atomicAdd(list_n, 1u);
    list[slot] = static_cast<uint32_t>(i);
    return;
  }
  x2m[i] = __double2float_rd(s * (1.0 - kLoMargin));
  cxb[i] = __double2float_ru(2.0 * kC1 * sqrt(s) * (1.0 + 9.5367431640625e-07));
  qa1[i] = __double2float_ru(2.0 * sqrt(s) * (1.0 + 9.5367431640625e-07));
  qa2[i] = __double2float_ru(2.0 * (sqrt(sr) + 128.0 * 2.384185791015625e-07 * sqrt(sh)) *
                             (1.0 + 9.5367431640625e-07));
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
  static constexpr int NP = (VEC + kNT - 1) / kNT;       // pieces per thread
  static constexpr int SLOT = kRows * RS * 2 + 4 * kRows * 4;  // rows | y2m | nyb | nry | nyh (bytes)
  static_assert(NP <= 2, "at most two pieces per thread");
};

// Dynamic LDS of the scan: eight tile slots | trans[kWaves][64][32].
constexpr int kSlots = 8;
template <int KS>
size_t scan_lds_bytes() {
  return kSlots * static_cast<size_t>(Scan<KS>::SLOT) + sizeof(float) * kWaves * 64 * 32;
}

// Candidate lists.  Per query, lists are laid out back to back: first
// `nrun` lists of `cap_run` entries (one per segment of a running launch),
// then lists of `cap_fix` entries written by fixed-threshold launches, two
// per segment (one per half of the tile's rows, each in row order; the
// replay merges the pair).  Lists follow the rows' order.
struct ListLayout {
  int nrun, cap_run;  // running lists
  int nfix, cap_fix;  // fixed-threshold lists (pairs)
  __host__ __device__ int nlists() const { return nrun + nfix; }
  __host__ __device__ int64_t per_query() const {
    return static_cast<int64_t>(nrun) * cap_run + static_cast<int64_t>(nfix) * cap_fix;
  }
  __host__ __device__ int64_t offset(int l) const {
    return l < nrun ? static_cast<int64_t>(l) * cap_run
                    : static_cast<int64_t>(nrun) * cap_run + static_cast<int64_t>(l - nrun) * cap_fix;
  }
  __host__ __device__ int cap(int l) const { return l < nrun ? cap_run : cap_fix; }
};

// One scan launch covers rows [row0, row_end) in nseg segments of seg_len
// rows.  A running launch (MODE 0) keeps each query's k smallest U, from
// empty, and writes segment s to list lbase + s.  A fixed launch (MODE 2)
// tests every row against the seed's k-th value and writes segment s to
// lists lbase + 2 s + half.  A seed holds, per query and ascending, k
// values each at or above the distance s of a distinct row before row0
// (exact distances of earlier candidates, rounded up to f32).
struct ScanArgs {
  const uint16_t* rows_b;
  const float* y2m;
  const float* nyb;
  const float* nry;
  int64_t npad;  // y2m, nyb, nry, nyh: consecutive arrays of npad floats
  int64_t n;
  const uint16_t* q_b;
  const float* x2m;
  const float* cxbq;
  const float* qa1;
  const float* qa2;
  int64_t q;
  int k;
  int64_t row0, row_end, seg_len;
  int nseg;
  int64_t nqt;
  ListLayout lay;
  int lbase;
  uint32_t* cand_idx;  // q x lay.per_query()
  uint32_t* cand_count;  // q x lay.nlists()
  const float* seed;   // q x k, or null
  const uint32_t* fold;  // (npad + kScanPadRows) x 4 folded-column dwords
};

// Running launch.  KT: registers of the per-query top-k (16 for k <= 16,
// else 32).
template <int KS, int KT>
__global__ __launch_bounds__(kNT, 2) void mixed_scan_kernel(const ScanArgs a) {
  using S = Scan<KS>;
  const uint16_t* __restrict__ rows_b = a.rows_b;
  const float* __restrict__ y2m = a.y2m;
  const int64_t n = a.n;
  const uint16_t* __restrict__ q_b = a.q_b;
  const float* __restrict__ x2m = a.x2m;
  const float* __restrict__ cxbq = a.cxbq;
  const int64_t q = a.q;
  const int k = a.k;
  const int nseg = a.nseg;
  const int64_t nqt = a.nqt;
  uint32_t* __restrict__ cand_idx = a.cand_idx;
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

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int col = lane & 31;   // query column inside a 32-wide block / row of A
  const int half = lane >> 5;  // K half / row quad; also the column block the lane owns
  const int64_t qw0 = qt * kQPB + wave * kQPW;
  const int64_t r_begin = a.row0 + static_cast<int64_t>(seg) * a.seg_len;
  const int64_t r_end = min(a.row_end, r_begin + a.seg_len);
  const int64_t ntiles = r_end > r_begin ? (r_end - r_begin + kRows - 1) / kRows : 0;
  const int nl = a.lay.nlists();
  const int64_t perq = a.lay.per_query();
  const int lseg = a.lbase + seg;  // this segment's list
  const int64_t loff = a.lay.offset(lseg);
  const int lcap = a.lay.cap(lseg);
  // rare-path transpose: each lane's 32 values of a tile (both column blocks)
  float* trans = reinterpret_cast<float*>(smem + kSlots * S::SLOT);

  // per column block values as two-element vectors, so that picking the
  // lane's own block is a register select, not an indexed private array
  bf16x8 bq[2][KS];
  f32x2 thr, x2v, cxq;
  i32x2 qvalid;
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
  // The lane owns query (column block `half`, column col) and keeps that
  // query's k smallest U in registers tk[0..k),
  // ascending, with tk[k..KT) held at -inf (the insertion below leaves them
  // there, so the k-th smallest is the maximum of tk and no register is
  // indexed by k), its candidate count and its threshold.
  const int64_t qo = qw0 + 32 * half + col;
  const double xo = static_cast<double>(half == 0 ? x2v.x : x2v.y);
  const double co = static_cast<double>(half == 0 ? cxq.x : cxq.y);
  const double a1o = qo < q ? static_cast<double>(a.qa1[qo]) : 0.0;
  const double a2o = qo < q ? static_cast<double>(a.qa2[qo]) : 0.0;
  float tk[KT];
#pragma unroll
  for (int j = 0; j < KT; ++j) tk[j] = j >= k ? -INFINITY : INFINITY;
  auto kth_of = [&]() {
    float m[KT];
#pragma unroll
    for (int j = 0; j < KT; ++j) m[j] = tk[j];
#pragma unroll
    for (int w = KT / 2; w >= 1; w /= 2)
#pragma unroll
      for (int j = 0; j < w; ++j) m[j] = fmaxf(m[j], m[j + w]);
    return m[0];
  };
  uint32_t cnt = 0;
  // both lanes of a column hold the thresholds of both column blocks
  auto set_thr = [&](float own) {
    const float other = __shfl_xor(own, 32);
    thr.x = half == 0 ? own : other;
    thr.y = half == 0 ? other : own;
  };

  // tile loads: thread tid carries pieces tid and tid + kNT (16 bytes of
  // rows each, when below VEC); threads 0..127 carry the tile's y2m (0..31),
  // nyb (32..63), nry (64..95) and nyh (96..127)
  struct TileRegs {
    u32x4 v[S::NP];
    float tv;
  };
  auto gload = [&](int64_t t, TileRegs& r) {
    const int64_t t0 = r_begin + t * kRows;
#pragma unroll
    for (int pc = 0; pc < S::NP; ++pc) {
      const int e = tid + kNT * pc;
      if (e < S::VEC) {
        const int64_t row = min<int64_t>(t0 + e / (S::LD / 8), n - 1);
        r.v[pc] = *reinterpret_cast<const u32x4*>(rows_b + row * S::LD + 8 * (e % (S::LD / 8)));
      }
    }
    // y2m, nyb, nry and nyh are consecutive arrays of a.npad floats
    if (tid < 128) r.tv = y2m[(tid >> 5) * a.npad + t0 + (tid & 31)];
  };
  auto swrite = [&](int64_t t, const TileRegs& r) {
    unsigned char* base = smem + static_cast<int>(t & (kSlots - 1)) * S::SLOT;
#pragma unroll
    for (int pc = 0; pc < S::NP; ++pc) {
      const int e = tid + kNT * pc;
      if (e < S::VEC)
        *reinterpret_cast<u32x4*>(base + 2 * ((e / (S::LD / 8)) * S::RS + 8 * (e % (S::LD / 8)))) =
            r.v[pc];
    }
    if (tid < 128) reinterpret_cast<float*>(base + 2 * kRows * S::RS)[tid] = r.tv;
  };

  // LDS holds eight tiles.  Every fourth step (t = 0 mod 4) stores tiles
  // t + 4 .. t + 7, loaded into registers four steps earlier, into the
  // slots of tiles t - 4 .. t - 1, then loads tiles t + 8 .. t + 11.  A
  // barrier after every fourth step (t = 3 mod 4) orders both the stores
  // after the last reads of those slots and the next reads after the
  // stores, so waves may drift apart by up to four steps, and the wait
  // before a store covers loads issued four steps earlier.
  auto slot_ptr = [&](int64_t t) { return smem + static_cast<int>(t & (kSlots - 1)) * S::SLOT; };
  auto mfma_tile = [&](int64_t t, f32x16 (&acc)[2]) {
    const uint16_t* A = reinterpret_cast<const uint16_t*>(slot_ptr(t));
    acc[0] = f32x16{};
    acc[1] = f32x16{};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const bf16x8 fa = load_frag(A + col * S::RS + 16 * ks + 8 * half);
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, bq[0][ks], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, bq[1][ks], acc[1], 0, 0, 0);
    }
  };
  // fast epilogue of tile t: acc <- L - x2m with the loose bound
  // (e = cxb nyb, one FMA and one add); returns per column block whether
  // any lane of the wave has a row below its query's threshold
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
      hit[cb] = __ballot(qvalid[cb] != 0 && m < thr[cb]) != 0;
    }
  };
  // Tight bound of one row: with x = bf16(x) + rx and y = bf16(y) + ry,
  // |x.y - C| <= ||x|| ||ry|| + (||rx|| + 128 * 2^-22 ||bf16(x)||) ||bf16(y)||
  // and ||bf16(y)|| <= nyh, so 2 |x.y - C| <= et = qa1 nry + qa2 nyh.
  auto tight_e = [&](double q1, double q2, float nr, float nh) {
    return (q1 * static_cast<double>(nr) + q2 * static_cast<double>(nh)) *
           (1.0 + 4.440892098500626e-16);
  };
  // Rare path of tile t (see the file comment): each lane handles its own
  // query, the rows of the tile below its threshold, in row order.
  auto slow = [&](int64_t t, f32x16 (&acc)[2]) {
    const unsigned char* base = slot_ptr(t);
    const float* ty2 = reinterpret_cast<const float*>(base + 2 * kRows * S::RS);
    const float* tny = ty2 + kRows;
    const float* trr = tny + kRows;
    const float* tnh = trr + kRows;
    const int64_t t0 = r_begin + t * kRows;
    uint32_t m16[2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      m16[cb] = 0;
#pragma unroll
      for (int v = 0; v < 16; ++v)
        m16[cb] |= static_cast<uint32_t>(qvalid[cb] != 0 && acc[cb][v] < thr[cb]) << v;
    }
    float* tw = trans + (wave * 64 + lane) * 32;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int g = 0; g < 4; ++g)
        *reinterpret_cast<f32x4*>(tw + 16 * cb + 4 * g) =
            f32x4{acc[cb][4 * g], acc[cb][4 * g + 1], acc[cb][4 * g + 2], acc[cb][4 * g + 3]};
    // lane value v of column block cb is row 8 (v / 4) + 4 half + v % 4:
    // the own lane has the rows of the own block with that half, the
    // partner lane the others
    const uint32_t mo = half == 0 ? m16[0] : m16[1];
    const uint32_t mp = __shfl_xor(half == 0 ? m16[1] : m16[0], 32);
    uint32_t M = 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const uint32_t o4 = (mo >> (4 * g)) & 0xFu, p4 = (mp >> (4 * g)) & 0xFu;
      M |= (half == 0 ? o4 : p4) << (8 * g);
      M |= (half == 0 ? p4 : o4) << (8 * g + 4);
    }
    float tcur = half == 0 ? thr.x : thr.y;
    float kcur = kth_of();
    while (M != 0) {
      const int r = __builtin_ctz(M);
      M &= M - 1;
      const int64_t row = t0 + r;
      if (row >= r_end) break;  // rows ascend: the rest are past the segment
      const int h = (r >> 2) & 1;
      const int v = (r >> 3) * 4 + (r & 3);
      const float val = trans[(wave * 64 + col + 32 * h) * 32 + 16 * half + v];
      if (!(val < tcur)) continue;  // the threshold fell earlier in this tile
      // the epilogue subtracted the loose e = cxb nyb; swap it for et
      const double et = tight_e(a1o, a2o, trr[r], tnh[r]);
      const double L = xo + static_cast<double>(val) + (co * static_cast<double>(tny[r]) - et);
      if (!(L < static_cast<double>(kcur))) continue;
      const double U = L + 2.0 * et + (kHiMargin + 2.0 * kLoMargin) * (xo + static_cast<double>(ty2[r]));
      const float uf = __double2float_ru(U);
      // insert into the ascending tk[0..k): slot j takes
      // min(tk[j], max(tk[j-1], uf)); slots past k stay -inf
#pragma unroll
      for (int j = KT - 1; j > 0; --j) tk[j] = fminf(tk[j], fmaxf(tk[j - 1], uf));
      tk[0] = fminf(tk[0], uf);
      if (cnt < static_cast<uint32_t>(lcap)) {
        const size_t sl = static_cast<size_t>(qo) * perq + loff + cnt;
        cand_idx[sl] = static_cast<uint32_t>(row);
      }
      ++cnt;
      // T - x2m rounded up: the f64 difference of two f32 values rounds
      // when their exponents are far apart, so it is taken rounded toward
      // +inf before the narrowing (also toward +inf)
      kcur = kth_of();
      tcur = __double2float_ru(__ocml_sub_rtp_f64(static_cast<double>(kcur), xo));
    }
    set_thr(tcur);
  };
  TileRegs tr[4];
  f32x16 acc[2];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < ntiles) gload(j, tr[j]);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < ntiles) swrite(j, tr[j]);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (4 + j < ntiles) gload(4 + j, tr[j]);
  __syncthreads();
  for (int64_t t = 0; t < ntiles; ++t) {
    mfma_tile(t, acc);
    bool hit[2];
    epilogue(t, acc, hit);
    if (hit[0] || hit[1]) slow(t, acc);
    if ((t & 3) == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (t + 4 + j < ntiles) swrite(t + 4 + j, tr[j]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (t + 8 + j < ntiles) gload(t + 8 + j, tr[j]);
    }
    if ((t & 3) == 3) __syncthreads();
  }
  if (qo < q) a.cand_count[static_cast<size_t>(qo) * nl + lseg] = cnt;
}

__device__ __forceinline__ uint16_t bf16_ru_pos(float f) {
  // f >= 0 finite: the smallest BF16 >= f
  const uint32_t u = __float_as_uint(f);
  return static_cast<uint16_t>((u >> 16) + ((u & 0xffffu) != 0u ? 1u : 0u));
}

// Fixed-threshold launch (see the file comment).  The bound's terms are
// folded into one more MFMA K step: a tile row carries, after its ld
// values, the BF16 columns [y2m hi, mid, lo, RU(nry), RU(nyh), 0 ...] and
// the query side [1, 1, 1, -RU(qa1), -RU(qa2), 0 ...], so the MFMA leaves
// L - x2m = y2m - 2C - e (e = RU(qa1) RU(nry) + RU(qa2) RU(nyh), exact BF16
// products, >= the tight bound) and the epilogue is a minimum.  The FP32
// roundings of those adds and of the y2m split stay inside the 2^-18
// margin.  The folded columns of every row are built once with the index
// (fold_terms_kernel).  Each lane writes its own rows (its half of the
// tile) of both column blocks' queries to the queries' lists of that half.
//
// Schedule.  Sixteen tile slots (152 KiB of LDS at d = 128; the registers
// allow one workgroup per CU, two waves per SIMD).  Step t issues the MFMAs
// of tile t + 1 with the minimum of tile t's results between them, so a
// wave keeps the matrix pipe fed during its own epilogue; the candidate
// path (rare) is a branch after the MFMAs.  Step t also stores tile t + 9
// from registers into its slot and loads tile t + 13 into the same
// registers (four tiles in flight per thread).  The slot of tile t + 9 last
// held tile t - 7, whose fragments are read in step t - 8, and the new tile
// is read in step t + 8.  A barrier at the start of every step 8 m falls in
// (t - 8, t] and in (t, t + 8]: one barrier per eight tiles.  Steps run in
// groups of eight (ntiles rounded up).  Tiles past the segment read real or
// padding rows (the BF16 rows and the folded columns carry kScanPadRows
// rows of padding past the index) and their rows fail the row < r_end test.
template <int KS>
struct FixTile {
  static constexpr int LD = 16 * KS;
  static constexpr int RS = LD + 16 + 8;      // row | folded K step | 16 B padding (bf16)
  static constexpr int VEC = kRows * LD / 8;  // 16-byte pieces of a tile's rows
  static constexpr int SLOT = kRows * RS * 2;  // bytes
  static_assert(VEC <= kNT, "one row piece per thread");
};
constexpr int kFixSlots = 16;
constexpr int kFixAhead = 9;  // step t stores tile t + kFixAhead
constexpr int kFixRing = 4;   // steps between a tile's load and its store
constexpr int kFixGroup = 8;  // steps per barrier
// the store of step t into the slot read last in step t + kFixAhead - kFixSlots - 1
// and next in step t + kFixAhead - 1: a barrier step in each interval
static_assert(kFixAhead - 1 >= kFixGroup && kFixSlots + 1 - kFixAhead >= kFixGroup,
              "a barrier between each slot's last read and its next store, and after it");
template <int KS>
size_t fixed_lds_bytes() {
  return kFixSlots * static_cast<size_t>(FixTile<KS>::SLOT);
}

// Folded columns of every row (four dwords): [hi | mid << 16, lo | RU(nry)
// << 16, RU(nyh), 0], hi + mid + lo = y2m up to 2^-24 y2m.  Rows from npad on
// are padding (y2m = +inf: a NaN sum that never passes a test).
__global__ void fold_terms_kernel(const float* __restrict__ y2m, const float* __restrict__ nry,
                                  const float* __restrict__ nyh, u32x4* __restrict__ fold,
                                  int64_t npad, int64_t nfold) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nfold) return;
  const float y = i < npad ? y2m[i] : INFINITY;
  const uint32_t h = bf16_rne(y);
  const float e1 = y - __uint_as_float(h << 16);
  const uint32_t m = bf16_rne(e1);
  const uint32_t l = bf16_rne(e1 - __uint_as_float(m << 16));
  const uint32_t c1 = i < npad ? bf16_ru_pos(nry[i]) : 0u;
  const uint32_t c2 = i < npad ? bf16_ru_pos(nyh[i]) : 0u;
  fold[i] = u32x4{h | (m << 16), l | (c1 << 16), c2, 0u};
}

template <int KS>
__global__ __launch_bounds__(kNT, 2) void fixed_scan_kernel(const ScanArgs a) {
  using S = FixTile<KS>;
  const uint16_t* __restrict__ rows_b = a.rows_b;
  const uint32_t* __restrict__ fold = a.fold;
  const int64_t q = a.q;
  const int k = a.k;
  const int nseg = a.nseg;
  extern __shared__ __align__(16) unsigned char smem[];
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
  if (qt >= a.nqt) return;  // whole workgroup leaves together

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int col = lane & 31;
  const int half = lane >> 5;
  const int64_t qw0 = qt * kQPB + wave * kQPW;
  const int64_t r_begin = a.row0 + static_cast<int64_t>(seg) * a.seg_len;
  const int64_t r_end = min(a.row_end, r_begin + a.seg_len);
  const int64_t ntiles = r_end > r_begin ? (r_end - r_begin + kRows - 1) / kRows : 0;
  const int nl = a.lay.nlists();
  const int64_t perq = a.lay.per_query();
  const int lseg = a.lbase + 2 * seg + half;  // this lane's half-list of the segment
  const int64_t loff = a.lay.offset(lseg);
  const int lcap = a.lay.cap(lseg);

  bf16x8 bq[2][KS + 1];
  float thr[2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int64_t qi = qw0 + 32 * cb + col;
    const bool valid = qi < q && a.x2m[qi < q ? qi : 0] < INFINITY;
    const int64_t qr = qi < q ? qi : 0;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      bq[cb][ks] = load_frag(a.q_b + qr * S::LD + 16 * ks + 8 * half);
    // the seed's k-th value T; the epilogue threshold is T - x2m rounded up
    // (the f64 difference rounded toward +inf, then narrowed toward +inf:
    // at or above the exact T - x2m even when the f64 difference rounds).
    // Queries that take no part get -inf, which no value is below.
    const float x2q = a.x2m[qr], tq = a.seed[qr * k + k - 1];
    thr[cb] = valid ? __double2float_ru(__ocml_sub_rtp_f64(static_cast<double>(tq),
                                                           static_cast<double>(x2q)))
                    : -INFINITY;
    const uint16_t b1 = bf16_ru_pos(a.qa1[qr]), b2 = bf16_ru_pos(a.qa2[qr]);
    const uint32_t one = 0x3f80u, n1 = 0x8000u | b1, n2 = 0x8000u | b2;
    const u32x4 fq = half == 0 ? u32x4{one | (one << 16), one | (n1 << 16), n2, 0u}
                               : u32x4{0u, 0u, 0u, 0u};
    bq[cb][KS] = __builtin_bit_cast(bf16x8, fq);
  }
  uint32_t cntf[2] = {0u, 0u};

  if (ntiles == 0) {  // the same for the whole workgroup
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int64_t qi = qw0 + 32 * cb + col;
      if (qi < q) a.cand_count[static_cast<size_t>(qi) * nl + lseg] = 0u;
    }
    return;
  }

  // The second half of each row's folded K step is zero on both sides;
  // the LDS there must hold finite values, and no tile store touches it.
  static_assert(kFixSlots * kRows == kNT, "one zero piece per thread");
  *reinterpret_cast<u32x4*>(smem + (tid / kRows) * S::SLOT +
                            2 * ((tid % kRows) * S::RS + S::LD + 8)) = u32x4{0u, 0u, 0u, 0u};

  // Tile loads: thread tid carries row piece tid (when the tile has that
  // many) and dword tid % 128 of the tile's folded columns.  Addresses are
  // a per-tile base (the same for the whole wave) plus a fixed lane offset.
  const int prow = (tid < S::VEC ? tid : 0) / (S::LD / 8);
  const int pcol = 8 * ((tid < S::VEC ? tid : 0) % (S::LD / 8));
  const int foff = tid & 127;  // dword of the tile's 32 x 4 folded dwords
  const uint32_t lds_piece = 2 * (prow * S::RS + pcol);
  const uint32_t lds_fold = 2 * ((foff >> 2) * S::RS + S::LD) + 4 * (foff & 3);
  struct TileRegs {
    u32x4 v;
    uint32_t f;
  };
  auto gload = [&](int64_t t, TileRegs& r) {
    const int64_t t0 = r_begin + t * kRows;
    if (S::VEC == kNT || tid < S::VEC)
      r.v = *reinterpret_cast<const u32x4*>(rows_b + t0 * S::LD + (prow * S::LD + pcol));
    r.f = fold[t0 * 4 + foff];
  };
  auto swrite = [&](int sl, const TileRegs& r) {
    unsigned char* base = smem + sl * S::SLOT;
    if (S::VEC == kNT || tid < S::VEC) *reinterpret_cast<u32x4*>(base + lds_piece) = r.v;
    *reinterpret_cast<uint32_t*>(base + lds_fold) = r.f;
  };
  auto frags = [&](int sl, bf16x8 (&fa)[KS + 1]) {
    const uint16_t* A = reinterpret_cast<const uint16_t*>(smem + sl * S::SLOT);
#pragma unroll
    for (int ks = 0; ks <= KS; ++ks) fa[ks] = load_frag(A + col * S::RS + 16 * ks + 8 * half);
  };
  auto mfmas = [&](const bf16x8 (&fa)[KS + 1], f32x16 (&acc)[2]) {
    acc[0] = f32x16{};
    acc[1] = f32x16{};
#pragma unroll
    for (int ks = 0; ks <= KS; ++ks) {
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[ks], bq[0][ks], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[ks], bq[1][ks], acc[1], 0, 0, 0);
    }
  };
  // hits of tile t: rows of this lane's half, ascending
  auto hits = [&](int64_t t, const f32x16 (&acc)[2], bool h0, bool h1) {
    const int64_t t0 = r_begin + t * kRows;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      if (!(cb == 0 ? h0 : h1)) continue;
      const int64_t qi = qw0 + 32 * cb + col;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        // val < thr (T - x2m rounded up) keeps every row with L < T; a row
        // that only the rounding lets in is a harmless extra candidate
        if (acc[cb][v] < thr[cb]) {
          const int64_t row = t0 + 8 * (v >> 2) + 4 * half + (v & 3);
          if (row < r_end) {
            if (cntf[cb] < static_cast<uint32_t>(lcap))
              a.cand_idx[static_cast<size_t>(qi) * perq + loff + cntf[cb]] =
                  static_cast<uint32_t>(row);
            ++cntf[cb];
          }
        }
      }
    }
  };

  TileRegs tr[kFixRing];
  f32x16 accA[2], accB[2];
  // tiles 0 .. kFixAhead - 1 into their slots, the next kFixRing into registers
  static_assert(kFixAhead <= 3 * kFixRing, "prologue in three batches");
#pragma unroll
  for (int g = 0; g < kFixAhead; g += kFixRing) {
#pragma unroll
    for (int j = 0; j < kFixRing; ++j)
      if (g + j < kFixAhead) gload(g + j, tr[j]);
#pragma unroll
    for (int j = 0; j < kFixRing; ++j)
      if (g + j < kFixAhead) swrite(g + j, tr[j]);
  }
#pragma unroll
  for (int j = 0; j < kFixRing; ++j) gload(kFixAhead + j, tr[j]);
  __syncthreads();
  {
    bf16x8 fa[KS + 1];
    frags(0, fa);
    mfmas(fa, accA);
  }
  // step t: epilogue of tile t (cur), MFMAs of tile t + 1 (nxt), ring
  // register set t % kFixRing
  auto step = [&](auto jc, int64_t t, f32x16 (&cur)[2], f32x16 (&nxt)[2]) {
    constexpr int R = decltype(jc)::value % kFixRing;
    bf16x8 fa[KS + 1];
    frags(static_cast<int>((t + 1) & (kFixSlots - 1)), fa);
    swrite(static_cast<int>((t + kFixAhead) & (kFixSlots - 1)), tr[R]);
    gload(t + kFixAhead + kFixRing, tr[R]);
    float m0 = cur[0][0], m1 = cur[1][0];
#pragma unroll
    for (int v = 1; v < 16; ++v) {
      m0 = fminf(m0, cur[0][v]);
      m1 = fminf(m1, cur[1][v]);
    }
    mfmas(fa, nxt);
    const bool h0 = __ballot(m0 < thr[0]) != 0;
    const bool h1 = __ballot(m1 < thr[1]) != 0;
    // LDS reads first, then the ring's stores and loads, then each MFMA
    // followed by two VALU instructions (the minimum, addresses, tests)
    __builtin_amdgcn_sched_group_barrier(0x100, KS + 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x200, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x020, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
#pragma unroll
    for (int i = 0; i < 2 * (KS + 1); ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
    }
    if (h0 || h1) hits(t, cur, h0, h1);
  };
  const int64_t nsteps = (ntiles + kFixGroup - 1) / kFixGroup * kFixGroup;
  for (int64_t t = 0; t < nsteps; t += kFixGroup) {
    __syncthreads();
    step(std::integral_constant<int, 0>{}, t + 0, accA, accB);
    step(std::integral_constant<int, 1>{}, t + 1, accB, accA);
    step(std::integral_constant<int, 2>{}, t + 2, accA, accB);
    step(std::integral_constant<int, 3>{}, t + 3, accB, accA);
    step(std::integral_constant<int, 4>{}, t + 4, accA, accB);
    step(std::integral_constant<int, 5>{}, t + 5, accB, accA);
    step(std::integral_constant<int, 6>{}, t + 6, accA, accB);
    step(std::integral_constant<int, 7>{}, t + 7, accB, accA);
  }
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int64_t qi = qw0 + 32 * cb + col;
    if (qi < q) a.cand_count[static_cast<size_t>(qi) * nl + lseg] = cntf[cb];
  }
}

// One wavefront per query for the sequential passes below; four per
// workgroup.  LDS traffic of a wave is ordered by the hardware; the fence
// keeps the compiler from moving it across the exchange points.
constexpr int kWavesPerBlock = 4;
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(__double_as_longlong(v)), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(__double_as_longlong(v) >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned long long>(static_cast<uint32_t>(lo)));
}
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// The k smallest of a seed (null: +inf) and the exact distances, rounded
// up to f32, of the entries in lists [l0, l0 + nl) (candidates of rows after
// the seed's rows, distinct): the next seed, ascending.  Entries past a
// list's cap were not stored; leaving them out keeps a valid seed.  Lanes
// 0 .. k-1 hold the current k smallest in ascending order; a value enters
// by one shift across the lanes.
__global__ __launch_bounds__(64 * kWavesPerBlock) void select_seed_kernel(
    const float* __restrict__ seed_in, const double* __restrict__ cand_S,
    const uint32_t* __restrict__ cand_count, ListLayout lay, int l0, int nl, int64_t q, int k,
    float* __restrict__ seed_out) {
  const int lane = threadIdx.x & 63;
  const int64_t qi = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (qi >= q) return;  // the whole wave
  float b = (lane < k && seed_in) ? seed_in[qi * k + lane] : INFINITY;
  if (lane >= k) b = INFINITY;
  float kth = lane_value(b, k - 1);
  const int64_t perq = lay.per_query();
  for (int l = l0; l < l0 + nl; ++l) {
    const uint32_t c = min(cand_count[static_cast<size_t>(qi) * lay.nlists() + l],
                           static_cast<uint32_t>(lay.cap(l)));
    const double* sv = cand_S + static_cast<size_t>(qi) * perq + lay.offset(l);
    for (uint32_t c0 = 0; c0 < c; c0 += 64) {
      const float v = c0 + lane < c ? __double2float_ru(sv[c0 + lane]) : INFINITY;
      uint64_t m = __ballot(v < kth);
      while (m) {
        const int src = __builtin_ctzll(m);
        const float x = lane_value(v, src);
        float prev = __shfl_up(b, 1);
        if (lane == 0) prev = -INFINITY;
        if (lane < k && x < b) b = x < prev ? prev : x;
        kth = lane_value(b, k - 1);
        m &= m - 1;
        m &= __ballot(v < kth);
      }
    }
  }
  if (lane < k) seed_out[qi * k + lane] = b;
}

// The reference heap (heap_select.rs) held across a wave's lanes: lane j < k
// holds entry j (distance hd, row hi).  Replacing the root and sifting it
// down follows HeapSelection::sift_down (heap_select.rs:71-86: the children
// of kk are 2kk and 2kk + 1; stop on Equal or Greater) with the node values
// read by readlane and the swap done by a select in every lane, so an
// insertion costs a few instructions per level and no memory traffic.
__device__ __forceinline__ void lane_heap_push_root(double& hd, uint32_t& hi, int lane, int k,
                                                    double d, uint32_t row) {
  if (lane == 0) {
    hd = d;
    hi = row;
  }
  const int n = k - 1;
  int kk = 0;
  double vk = d;
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

// Pass 2a: the exact s of every stored candidate (the reference's
// arithmetic: f32 differences and squares, f64 sum in dimension order).
// One workgroup per query; its threads take the candidates in turn.
constexpr int kRescoreThreads = 128;
__global__ __launch_bounds__(kRescoreThreads) void mixed_rescore_kernel(
    const float* __restrict__ rows, const float* __restrict__ qs, int ld, int64_t q, ListLayout lay,
    int l_lo, int l_hi, const uint32_t* __restrict__ cand_idx,
    const uint32_t* __restrict__ cand_count, const float* __restrict__ x2m,
    double* __restrict__ cand_S) {
  const int64_t qi = blockIdx.x;
  if (qi >= q || !(x2m[qi] < INFINITY)) return;  // uniform: listed for the exact kernel
  __shared__ float qrow[16 * kMaxKS];
  __shared__ uint32_t start[kMaxLists + 1];
  const int nl = lay.nlists();
  for (int j = threadIdx.x; j < ld; j += kRescoreThreads) qrow[j] = qs[qi * ld + j];
  const int nr = l_hi - l_lo;  // lists l_lo .. l_hi - 1
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < nr; ++i) {
      start[i] = acc;
      acc += min(cand_count[static_cast<size_t>(qi) * nl + l_lo + i],
                 static_cast<uint32_t>(lay.cap(l_lo + i)));
    }
    start[nr] = acc;
  }
  __syncthreads();
  const size_t qbase = static_cast<size_t>(qi) * lay.per_query();
  for (uint32_t e = threadIdx.x; e < start[nr]; e += kRescoreThreads) {
    int i = 0;
    while (start[i + 1] <= e) ++i;
    const size_t slot = qbase + lay.offset(l_lo + i) + (e - start[i]);
    // the row read 16 values at a time (ld is a multiple of 16, rows are
    // 64-byte aligned), the terms added in dimension order
    const f32x4* r4 = reinterpret_cast<const f32x4*>(rows + static_cast<int64_t>(cand_idx[slot]) * ld);
    double sum = 0.0;
    for (int j = 0; j < ld; j += 16) {
      f32x4 v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = r4[j / 4 + i];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          sum = sum + Term<float, kEuclid>::f(qrow[j + 4 * i + c], v[i][c], 2.0);
    }
    cand_S[slot] = sum;
  }
}

// Pass 2b: replay of the heap rule over the candidates in row order, one
// wavefront per query.  Entries are taken 64 at a time (lane = entry); the
// lanes whose distance is below the heap's largest (the reference's strict
// <) enter one after another, in row order, by a sift-down of the heap
// held across the lanes, and the rest are tested again against the new top.  A
// fixed launch's list pair is merged by row 64 entries at a time: an entry's
// place is its lane plus the count of smaller rows in the other list's
// window, and the places below 64 are the next 64 rows of the merge.
// Queries whose lists overflowed go to the serial exact kernel's list.
__global__ __launch_bounds__(64 * kWavesPerBlock) void mixed_replay_kernel(
    int64_t q, int k, ListLayout lay, const uint32_t* __restrict__ cand_idx,
    const double* __restrict__ cand_S, const uint32_t* __restrict__ cand_count,
    const float* __restrict__ x2m, int64_t* __restrict__ out_idx, double* __restrict__ out_dist,
    uint32_t* __restrict__ out_count, uint32_t* __restrict__ list, uint32_t* __restrict__ list_n,
    unsigned long long* __restrict__ counters) {
  __shared__ uint32_t s_win[kWavesPerBlock][2][64];
  __shared__ uint32_t s_mrow[kWavesPerBlock][64];
  __shared__ double s_mS[kWavesPerBlock][64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t qi = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  if (qi >= q) return;                     // the whole wave
  if (!(x2m[qi] < INFINITY)) return;       // listed for the exact kernel already
  const int nl = lay.nlists();
  const uint32_t* cc = cand_count + static_cast<size_t>(qi) * nl;
  bool over = false;
  for (int l = lane; l < nl; l += 64) over = over || cc[l] > static_cast<uint32_t>(lay.cap(l));
  if (__ballot(over) != 0) {
    if (lane == 0) {
      const uint32_t slot = atomicAdd(list_n, 1u);
      list[slot] = static_cast<uint32_t>(qi);
    }
    return;
  }
  // the heap in lanes 0 .. k-1 (the reference seeds it with k x {inf, None})
  double hd = INFINITY;
  uint32_t hi = 0xffffffffu;
  double top = INFINITY;
  // entries of one chunk, in row order over the lanes (valid lanes first)
  auto consume = [&](double S, uint32_t row, bool valid) {
    const double d = __builtin_sqrt(S);
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
  };
  const size_t qbase = static_cast<size_t>(qi) * lay.per_query();
  unsigned long long seen = 0;
  for (int l = 0; l < lay.nrun; ++l) {
    const size_t base = qbase + lay.offset(l);
    const uint32_t n = cc[l];
    seen += n;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
      const bool valid = c0 + lane < n;
      const double S = valid ? cand_S[base + c0 + lane] : 0.0;
      const uint32_t row = valid ? cand_idx[base + c0 + lane] : 0u;
      consume(S, row, valid);
    }
  }
  uint32_t* winA = s_win[w][0];
  uint32_t* winB = s_win[w][1];
  uint32_t* mrow = s_mrow[w];
  double* mS = s_mS[w];
  // count of entries below v in a window of 64 ascending rows (absent
  // entries are 0xffffffff)
  auto rank_in = [&](const uint32_t* win, uint32_t v) {
    int lo = 0;
#pragma unroll
    for (int step = 64; step > 0; step >>= 1)
      if (lo + step <= 64 && win[lo + step - 1] < v) lo += step;
    return lo;
  };
  for (int l = lay.nrun; l < nl; l += 2) {
    const size_t bA = qbase + lay.offset(l), bB = qbase + lay.offset(l + 1);
    const uint32_t nA = cc[l], nB = cc[l + 1];
    seen += nA + nB;
    uint32_t ia = 0, ib = 0;
    while (ia < nA || ib < nB) {
      const bool okA = ia + lane < nA, okB = ib + lane < nB;
      const uint32_t ra = okA ? cand_idx[bA + ia + lane] : 0xffffffffu;
      const uint32_t rb = okB ? cand_idx[bB + ib + lane] : 0xffffffffu;
      const double Sa = okA ? cand_S[bA + ia + lane] : 0.0;
      const double Sb = okB ? cand_S[bB + ib + lane] : 0.0;
      winA[lane] = ra;
      winB[lane] = rb;
      wave_sync();
      const int pa = lane + rank_in(winB, ra);
      const int pb = lane + rank_in(winA, rb);
      const bool takeA = okA && pa < 64, takeB = okB && pb < 64;
      if (takeA) {
        mrow[pa] = ra;
        mS[pa] = Sa;
      }
      if (takeB) {
        mrow[pb] = rb;
        mS[pb] = Sb;
      }
      const uint32_t cA = static_cast<uint32_t>(__popcll(__ballot(takeA)));
      const uint32_t cB = static_cast<uint32_t>(__popcll(__ballot(takeB)));
      wave_sync();
      const bool valid = static_cast<uint32_t>(lane) < cA + cB;
      const double S = valid ? mS[lane] : 0.0;
      const uint32_t row = valid ? mrow[lane] : 0u;
      wave_sync();  // the window and merge slots are rewritten next round
      consume(S, row, valid);
      ia += cA;
      ib += cB;
    }
  }
  if (lane == 0) {
    atomicAdd(&counters[0], seen);
    atomicAdd(&counters[1], seen);
  }
  // the heap array without its empty slots, in heap order
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
  // BF16 rows, then kScanPadRows zero rows; terms: y2m | nyb | nry | nyh
  // (npad floats each) | folded columns (npad + kScanPadRows rows of 4 dwords)
  const int64_t nfold = npad + kScanPadRows;
  SC_HIP(hipMalloc(&b, static_cast<size_t>(n + kScanPadRows) * ldb * sizeof(uint16_t)));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&terms), sizeof(float) * (4 * npad + 4 * nfold)));
  SC_HIP(hipMemsetAsync(static_cast<uint16_t*>(b) + static_cast<size_t>(n) * ldb, 0,
                        static_cast<size_t>(kScanPadRows) * ldb * sizeof(uint16_t), s));
  DevBuf flag;
  SC_HIP(flag.alloc(4, s));
  SC_HIP(hipMemsetAsync(flag.p, 0, 4, s));
  hipLaunchKernelGGL(fill_terms_kernel, dim3(static_cast<unsigned>((npad - n + 255) / 256)), dim3(256),
                     0, s, terms, terms + npad, terms + 2 * npad, terms + 3 * npad, n, npad);
  if (n > 0)
    hipLaunchKernelGGL(prep_rows_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       s, static_cast<const float*>(ix->rows), n, static_cast<int>(ix->ld),
                       static_cast<uint16_t*>(b), terms, terms + npad, terms + 2 * npad,
                       terms + 3 * npad, flag.as<uint32_t>());
  hipLaunchKernelGGL(fold_terms_kernel, dim3(static_cast<unsigned>((nfold + 255) / 256)), dim3(256), 0,
                     s, terms, terms + 2 * npad, terms + 3 * npad,
                     reinterpret_cast<u32x4*>(terms + 4 * npad), npad, nfold);
  SC_HIP(hipGetLastError());
  uint32_t bad = 0;
  SC_HIP(hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  ix->rows_bf16 = b;
  ix->norms2 = terms;  // y2m | nyb | nry | nyh (npad floats each) | folded columns
  ix->ldb = bad ? 0 : static_cast<size_t>(ldb);
  *bad_out = bad != 0;
  return SC_OK;
}

template <int KS, int KT>
void launch_running(dim3 grid, hipStream_t stream, const ScanArgs& a) {
  const size_t lds = scan_lds_bytes<KS>();
  static const bool attr = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&mixed_scan_kernel<KS, KT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(scan_lds_bytes<KS>()));
    return true;
  }();
  (void)attr;
  hipLaunchKernelGGL((mixed_scan_kernel<KS, KT>), grid, dim3(kNT), lds, stream, a);
}

template <int KS>
void launch_fixed(dim3 grid, hipStream_t stream, const ScanArgs& a) {
  const size_t lds = fixed_lds_bytes<KS>();
  static const bool attr = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&fixed_scan_kernel<KS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(fixed_lds_bytes<KS>()));
    return true;
  }();
  (void)attr;
  hipLaunchKernelGGL((fixed_scan_kernel<KS>), grid, dim3(kNT), lds, stream, a);
}

// mode: 0 running, 2 fixed threshold (from a seed)
int launch_scan(int ks, int mode, const ScanArgs& a, hipStream_t stream) {
  const int64_t nblocks = a.nseg <= kXcds
                              ? (a.nqt + kXcds / a.nseg - 1) / (kXcds / a.nseg) * kXcds
                              : a.nqt * a.nseg;
  const dim3 grid(static_cast<unsigned>(nblocks));
  switch (ks) {
#define SC_SCAN_CASE(KS)                                      \
  case KS:                                                   \
    if (mode == 2)                                           \
      launch_fixed<KS>(grid, stream, a);                     \
    else if (a.k <= 16)                                      \
      launch_running<KS, 16>(grid, stream, a);               \
    else                                                     \
      launch_running<KS, 32>(grid, stream, a);               \
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

  // plan: ~kTargetBlocks workgroups (one per CU).  Segments are a power of
  // two: up to 8 whole XCDs share a segment; small batches split the rows
  // further so that they still fill the chip.
  const int64_t nqt = (q + kQPB - 1) / kQPB;
  const int64_t tiles = std::max<int64_t>(1, (n + kRows - 1) / kRows);
  int nseg = 1;
  while (nseg < 64 && nqt * nseg < kTargetBlocks && nseg * 2 <= tiles) nseg *= 2;
  // Levels.  Small indexes: one running launch over all rows.  Indexes of
  // kSeedMinRows rows or more: a running launch over the first kPrefixRows
  // rows, then fixed-threshold launches over row ranges growing kLevelRatio
  // times.  The candidates of each level are rescored as it ends, and the
  // next level tests against the k-th smallest exact distance among all
  // earlier candidates (rounded up to f32).
  struct Level {
    int64_t r0, r1;
  };
  std::vector<Level> levels;
  if (n >= kSeedMinRows) {
    levels.push_back({0, kPrefixRows});
    for (int64_t b = kPrefixRows; b < n;) {
      const int64_t e = std::min<int64_t>(n, b * kLevelRatio);
      levels.push_back({b, e});
      b = e;
    }
  } else {
    levels.push_back({0, n});
  }
  auto seg_len_of = [&](const Level& L) {
    const int64_t lt = std::max<int64_t>(1, (L.r1 - L.r0 + kRows - 1) / kRows);
    return (lt + nseg - 1) / nseg * kRows;
  };
  const int nfix_levels = static_cast<int>(levels.size()) - 1;
  ListLayout lay{};
  lay.nrun = nseg;
  {
    const int64_t sl = seg_len_of(levels[0]);
    const double Lseg = static_cast<double>(std::min<int64_t>(sl, std::max<int64_t>(n, 1)));
    int64_t c = static_cast<int64_t>(4.0 * k * (1.0 + std::log(std::max(1.0, Lseg / k)))) + 128;
    lay.cap_run = static_cast<int>(std::min<int64_t>(c, sl));
  }
  lay.nfix = 2 * nseg * nfix_levels;
  lay.cap_fix = 0;
  for (int i = 1; i <= nfix_levels; ++i) {
    // expected entries of one half-list: rows of a segment times k over
    // the rows before the level, doubled for the bound's band
    const Level& L = levels[i];
    const double per_seg = static_cast<double>(L.r1 - L.r0) / nseg;
    const double expect = per_seg * k / static_cast<double>(L.r0);
    const int64_t c = static_cast<int64_t>(6.0 * expect) + 128;
    lay.cap_fix = static_cast<int>(std::max<int64_t>(lay.cap_fix, std::min<int64_t>(c, seg_len_of(L))));
  }
  const int64_t perq = lay.per_query();

  DevBuf qb, x2, cx, cidx, cS, ccnt, lst, flags, seedA, seedB;
  SC_HIP(qb.alloc(sizeof(uint16_t) * q * ld, stream));
  SC_HIP(x2.alloc(sizeof(float) * q, stream));
  SC_HIP(cx.alloc(sizeof(float) * 3 * q, stream));  // cxb | qa1 | qa2
  SC_HIP(cidx.alloc(sizeof(uint32_t) * q * perq, stream));
  SC_HIP(cS.alloc(sizeof(double) * q * perq, stream));
  SC_HIP(ccnt.alloc(sizeof(uint32_t) * q * lay.nlists(), stream));
  if (nfix_levels > 0) {
    SC_HIP(seedA.alloc(sizeof(float) * q * k, stream));
    SC_HIP(seedB.alloc(sizeof(float) * q * k, stream));
  }
  SC_HIP(lst.alloc(sizeof(uint32_t) * q, stream));
  SC_HIP(flags.alloc(64, stream));
  SC_HIP(hipMemsetAsync(flags.p, 0, 64, stream));
  uint32_t* list_n = flags.as<uint32_t>();
  unsigned long long* counters = reinterpret_cast<unsigned long long*>(flags.as<char>() + 8);
  const float* qf = static_cast<const float*>(d_queries);

  hipLaunchKernelGGL(prep_queries_kernel, dim3(static_cast<unsigned>((q + 255) / 256)), dim3(256),
                     0, stream, qf, q, ld, qb.as<uint16_t>(), x2.as<float>(), cx.as<float>(),
                     cx.as<float>() + q, cx.as<float>() + 2 * q, lst.as<uint32_t>(), list_n);
  SC_HIP(hipGetLastError());
  ScanArgs sa{};
  sa.rows_b = static_cast<const uint16_t*>(ix->rows_bf16);
  sa.y2m = y2m;
  sa.nyb = nyb;
  sa.nry = ix->norms2 + 2 * npad;
  sa.npad = npad;
  sa.fold = reinterpret_cast<const uint32_t*>(ix->norms2 + 4 * npad);
  sa.n = n;
  sa.q_b = qb.as<uint16_t>();
  sa.x2m = x2.as<float>();
  sa.cxbq = cx.as<float>();
  sa.qa1 = cx.as<float>() + q;
  sa.qa2 = cx.as<float>() + 2 * q;
  sa.q = q;
  sa.k = k;
  sa.nqt = nqt;
  sa.nseg = nseg;
  sa.lay = lay;
  sa.cand_idx = cidx.as<uint32_t>();
  sa.cand_count = ccnt.as<uint32_t>();
  float* seed_cur = seedA.as<float>();
  float* seed_next = seedB.as<float>();
  // exact distances of the candidates in lists [l_lo, l_hi)
  auto rescore = [&](int l_lo, int l_hi) -> int {
    hipLaunchKernelGGL(mixed_rescore_kernel, dim3(static_cast<unsigned>(q)), dim3(kRescoreThreads),
                       0, stream, static_cast<const float*>(ix->rows), qf, ld, q, lay, l_lo, l_hi,
                       cidx.as<uint32_t>(), ccnt.as<uint32_t>(), x2.as<float>(), cS.as<double>());
    SC_HIP(hipGetLastError());
    return SC_OK;
  };
  // the next seed from the current one (null: none yet) and lists [l0, l0 + nl)
  auto select = [&](const float* in, int l0, int nl) -> int {
    hipLaunchKernelGGL(select_seed_kernel,
                       dim3(static_cast<unsigned>((q + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(64 * kWavesPerBlock), 0, stream, in, cS.as<double>(), ccnt.as<uint32_t>(), lay, l0, nl, q, k,
                       seed_next);
    SC_HIP(hipGetLastError());
    std::swap(seed_cur, seed_next);
    return SC_OK;
  };
  {
    // every launch before the last level's scan is timed as the prefix
    ProfScope ps(nfix_levels > 0 ? "knn_mixed_prefix" : "knn_mixed_gemm", stream);
    ScanArgs la = sa;
    la.row0 = 0;
    la.row_end = levels[0].r1;
    la.seg_len = seg_len_of(levels[0]);
    la.lbase = 0;
    if (int rc = launch_scan(ld / 16, 0, la, stream)) return rc;
    if (nfix_levels > 0) {
      if (int rc = rescore(0, nseg)) return rc;
      if (int rc = select(nullptr, 0, nseg)) return rc;
    }
    for (int i = 1; i < nfix_levels; ++i) {
      ScanArgs fa = sa;
      fa.row0 = levels[i].r0;
      fa.row_end = levels[i].r1;
      fa.seg_len = seg_len_of(levels[i]);
      fa.lbase = nseg + 2 * nseg * (i - 1);
      fa.seed = seed_cur;
      if (int rc = launch_scan(ld / 16, 2, fa, stream)) return rc;
      if (int rc = rescore(fa.lbase, fa.lbase + 2 * nseg)) return rc;
      if (int rc = select(seed_cur, fa.lbase, 2 * nseg)) return rc;
    }
  }
  int l_last = 0;  // first list of the last scan
  if (nfix_levels > 0) {
    ProfScope ps("knn_mixed_gemm", stream);
    const int i = nfix_levels;
    ScanArgs fa = sa;
    fa.row0 = levels[i].r0;
    fa.row_end = levels[i].r1;
    fa.seg_len = seg_len_of(levels[i]);
    fa.lbase = nseg + 2 * nseg * (i - 1);
    fa.seed = seed_cur;
    l_last = fa.lbase;
    if (int rc = launch_scan(ld / 16, 2, fa, stream)) return rc;
  }
  {
    ProfScope ps("knn_mixed_rescore", stream);
    if (int rc = rescore(l_last, lay.nlists())) return rc;
    hipLaunchKernelGGL(mixed_replay_kernel,
                       dim3(static_cast<unsigned>((q + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(64 * kWavesPerBlock), 0, stream, q, k, lay, cidx.as<uint32_t>(),
                       cS.as<double>(), ccnt.as<uint32_t>(), x2.as<float>(), d_idx, d_dist, d_count,
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
