// One Lloyd step (BBDTree::clustering's outputs, FP64 form) that allocates
// nothing: for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// The step's scratch (the problem struct, the loop state and arrival
// counter, the reduce work lists, the padded centroids, the per-block and
// per-group partial sums and counts, the per-block ||x - mu||^2, the
// per-cluster distortion terms and the shift mu) lives in one file-scope device array,
// `scratch`, of kNoAllocScratchBytes.  Host code never passes the array's
// name to a kernel: the name is a host-side stand-in, not a device address.
// The first call on each device looks the array's address up with
// hipGetSymbolAddress (with that device current) and keeps it per device.
//
// Every call on one GPU shares that GPU's array.  Calls that overlap on one
// GPU (two host threads, or two streams whose work is in flight together)
// race on it; the caller serialises them.  Everything that must start at
// zero (the arrival counter, the loop state) is cleared by a kernel of the
// call itself, so no memset node is recorded when a caller captures it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <mutex>

#include "common.hpp"
#include "kmeans.hpp"

// 64 MiB per device.  The step needs about
//   8 (nb + ng) k d + 12 nb k + 8 k ld + small
// bytes, nb = ceil(n / 2048) and ng = ceil(nb / 16); at k = d = 64 that is
// about 17 MiB per million rows, so up to roughly 3.8 million rows.
constexpr size_t kNoAllocScratchBytes = size_t{64} << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace sc {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<unsigned char*, kMaxDevices> g_scratch_ptr{};

// The device address of this device's copy of `scratch` (looked up once).
int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(SC_ERR_UNSUPPORTED, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    SC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(SC_ERR_DEVICE, "scratch array has no device address");
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return SC_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Writes the problem struct and the reduce work lists of one problem into
// scratch and clears the arrival counter.
__global__ void noalloc_setup_kernel(KmProb p, KmProb* dst, uint2* group, int64_t ngroup,
                                     uint2* fin, int k) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t == 0) {
    *dst = p;
    *p.arrive = 0u;
  }
  for (int64_t e = t; e < ngroup; e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    group[e] = make_uint2(0u, static_cast<unsigned>(e));
  for (int64_t e = t; e < k; e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    fin[e] = make_uint2(0u, static_cast<unsigned>(e));
}

__global__ void noalloc_distortion_kernel(const KmState* st, double* out) { *out = st->dist; }

}  // namespace
}  // namespace sc

using namespace sc;

extern "C" int sc_kmeans_step_device_noalloc(const sc_index* index, const double* d_centroids,
                                             size_t k, double* d_sums, uint64_t* d_counts,
                                             uint32_t* d_membership, double* d_distortion,
                                             void* stream) {
  clear_error();
  if (!index || !d_centroids || !d_sums || !d_counts || !d_membership || !d_distortion)
    return fail(SC_PARAMETERS_ERROR, "null argument");
  if (k < 1) return fail(SC_PARAMETERS_ERROR, "k must be >= 1");
  if (k > (1u << 20)) return fail(SC_ERR_UNSUPPORTED, "k above 1048576");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = static_cast<int64_t>(index->n);
  const int64_t d = static_cast<int64_t>(index->d);
  const int64_t d1 = std::max<int64_t>(d, 1);
  const int64_t ld = static_cast<int64_t>(index->ld);
  const int64_t kk = static_cast<int64_t>(k);
  const int64_t nb = (n + kBlockRows - 1) / kBlockRows;
  const int64_t ng = (nb + kGroupBlocks - 1) / kGroupBlocks;
  const int64_t nb1 = std::max<int64_t>(nb, 1), ng1 = std::max<int64_t>(ng, 1);
  const int64_t slices = (kk * d + kReduceChunk - 1) / kReduceChunk;
  const int64_t ngroup = ng * slices;

  // scratch layout, every region 256-byte aligned
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t at = off;
    off += align256(bytes ? bytes : 1);
    return at;
  };
  const size_t o_prob = take(sizeof(KmProb));
  const size_t o_state = take(sizeof(KmState) + 16);
  const size_t o_group = take(sizeof(uint2) * static_cast<size_t>(std::max<int64_t>(ngroup, 1)));
  const size_t o_fin = take(sizeof(uint2) * static_cast<size_t>(kk));
  const size_t o_cent = take(sizeof(double) * static_cast<size_t>(kk * ld));
  const size_t o_psum = take(sizeof(double) * static_cast<size_t>(nb1 * kk * d1));
  const size_t o_pcount = take(sizeof(uint32_t) * static_cast<size_t>(nb1 * kk));
  const size_t o_gsum = take(sizeof(double) * static_cast<size_t>(ng1 * kk * d1));
  const size_t o_gcount = take(sizeof(uint64_t) * static_cast<size_t>(ng1 * kk));
  const size_t o_x2 = take(sizeof(double) * static_cast<size_t>(nb1 + 2 * kk + ld));
  if (off > kNoAllocScratchBytes)
    return fail(SC_ERR_UNSUPPORTED,
                "k-means step needs " + std::to_string(off) + " bytes of scratch; the no-alloc "
                "entry point has " + std::to_string(kNoAllocScratchBytes) +
                " (use sc_kmeans_step for this size)");

  SC_HIP(hipSetDevice(index->device));
  unsigned char* base = nullptr;
  if (int rc = scratch_address(index->device, &base)) return rc;

  // The launch helpers take a KmWork; its buffers stay empty (nothing is
  // allocated or freed) and every pointer goes into scratch or the caller's
  // outputs.
  KmWork w;
  KmProb& p = w.p;
  p = KmProb{};
  p.n = n;
  p.d = static_cast<int>(d);
  p.ld = static_cast<int>(ld);
  p.k = static_cast<int>(kk);
  p.kpad = (p.k + 31) / 32 * 32;
  p.nb = nb;
  p.ng = ng;
  p.rows = index->rows;
  p.cent = reinterpret_cast<double*>(base + o_cent);
  p.labels = d_membership;
  p.psum = reinterpret_cast<double*>(base + o_psum);
  p.pcount = reinterpret_cast<uint32_t*>(base + o_pcount);
  p.gsum = reinterpret_cast<double*>(base + o_gsum);
  p.gcount = reinterpret_cast<uint64_t*>(base + o_gcount);
  p.sums = d_sums;
  p.counts = d_counts;
  p.x2block = reinterpret_cast<double*>(base + o_x2);
  p.term = p.x2block + nb1;
  p.mu = p.term + 2 * kk;
  p.st = reinterpret_cast<KmState*>(base + o_state);
  p.arrive = reinterpret_cast<uint32_t*>(base + o_state + sizeof(KmState));
  w.dprob = reinterpret_cast<KmProb*>(base + o_prob);
  w.items.group = reinterpret_cast<uint2*>(base + o_group);
  w.items.fin = reinterpret_cast<uint2*>(base + o_fin);
  w.items.ngroup = ngroup;
  w.items.nfinal = kk;
  w.items.nprob = 1;

  const unsigned setup_grid =
      static_cast<unsigned>(std::min<int64_t>((std::max(ngroup, kk) + 255) / 256, 1024));
  hipLaunchKernelGGL(noalloc_setup_kernel, dim3(std::max(setup_grid, 1u)), dim3(256), 0, s, p,
                     w.dprob, w.items.group, ngroup, w.items.fin, p.k);
  SC_HIP(hipGetLastError());
  if (int rc = copy_padded_device(d_centroids, SC_F64, k, index->d, index->ld, p.cent, s))
    return rc;
  if (int rc = kmeans_state_init(p.st, s)) return rc;
  if (int rc = kmeans_x2_launch(index, w, s)) return rc;
  if (int rc = kmeans_assign_launch(index, index->rows, n, p.cent, p.k, p.labels, p.st, s))
    return rc;
  if (int rc = kmeans_partial_launch(index, w, p.k, p.st, s)) return rc;
  if (int rc = kmeans_reduce_launch(w.dprob, w.items, false, false, s)) return rc;
  hipLaunchKernelGGL(noalloc_distortion_kernel, dim3(1), dim3(1), 0, s, p.st, d_distortion);
  SC_HIP(hipGetLastError());
  return SC_OK;
}
